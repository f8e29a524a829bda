"""Python binding of the MI355X string sort (libgss.so, C ABI in include/gss.h).

Mirrors the reference's routine registry (src/routine.h:32-59,
src/routines.c:31-75): each registered routine has a name, a description, a
multicore flag and ``f(strings)``, which sorts a list of byte strings through
the routine's ``void f(unsigned char**, size_t)`` entry point.  Device-resident
sorting of torch tensors goes through :func:`sort_device`.

The HIP library is required: importing works without it, but every call that
needs it raises :class:`GssError` naming the missing file.  There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libgss.so"

GSS_OK = 0
GSS_MAX_PASSES = 64


class GssError(RuntimeError):
    pass


class Tuning(C.Structure):
    _fields_ = [("local_threshold", C.c_uint32), ("radix_bits", C.c_uint32), ("tile", C.c_uint32),
                ("key_bytes", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("text_bytes", C.c_uint64),
        ("global_passes", C.c_uint32),
        ("local_segments", C.c_uint32),
        ("active", C.c_uint64 * GSS_MAX_PASSES),
        ("local_strings", C.c_uint64),
        ("local_rounds_max", C.c_uint32),
        ("local_round_strings", C.c_uint64),
        ("local_level_strings", C.c_uint64),
        ("first_pass_fused", C.c_uint32),
    ]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("sample", C.c_uint64), ("signature", C.c_uint32), ("candidates", C.c_uint32)]


class PointerInfo(C.Structure):
    _fields_ = [("uploaded_text_bytes", C.c_uint64), ("pieces", C.c_uint64),
                ("packed", C.c_uint32), ("pad", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32),
                ("max_batch_requests", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("batches", C.c_uint64),
        ("strings", C.c_uint64),
        ("batch_capacity", C.c_uint64),
        ("max_batch_requests_seen", C.c_uint32),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("max_ms", C.c_double),
    ]


# Every symbol include/gss.h declares, with its ctypes signature.
_SIGNATURES = {
    "gss_msd": (None, [C.POINTER(C.c_void_p), C.c_size_t]),
    "gss_msd_lcp": (None, [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(C.c_size_t)]),
    "gss_set_pointer_limits": (C.c_int, [C.c_uint64, C.c_uint64]),
    "gss_pointer_last_info": (C.c_int, [C.POINTER(PointerInfo)]),
    "gss_sort_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_uint64, C.c_void_p]),
    "gss_sort_batch_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.POINTER(C.c_uint64), C.c_uint32,
                                          C.c_void_p]),
    "gss_ctx_create": (C.c_void_p, []),
    "gss_ctx_destroy": (None, [C.c_void_p]),
    "gss_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "gss_ctx_sort_batch_device32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64),
                                              C.c_uint32, C.c_void_p]),
    "gss_ctx_get_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gss_ctx_reserve": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gss_server_create": (C.c_void_p, [C.c_void_p, C.c_uint64, C.POINTER(ServerConfig)]),
    "gss_server_submit": (C.c_int64, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.c_void_p]),
    "gss_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "gss_server_done": (C.c_int, [C.c_void_p, C.c_int64]),
    "gss_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "gss_server_destroy": (None, [C.c_void_p]),
    "gss_ingest_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.c_void_p]),
    "gss_suffixes_device32": (C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p]),
    "gss_write_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                     C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "gss_merge_lcp_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                         C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "gss_merge_lcp_device32_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p,
                                                 C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                                 C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gss_merge_lcp_kway_device32": (C.c_int, [C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                              C.POINTER(C.c_uint64), C.c_uint32, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "gss_autotune_device32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                        C.POINTER(AutotuneResult), C.c_void_p]),
    "gss_autotune_enable": (C.c_int, [C.c_int]),
    "gss_autotune_reset": (C.c_int, []),
    "gss_autotune_cached": (C.c_int, []),
    "gss_get_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "gss_set_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "gss_get_stats": (C.c_int, [C.POINTER(Stats)]),
    "gss_profile_enable": (C.c_int, [C.c_int]),
    "gss_profile_reset": (C.c_int, []),
    "gss_profile_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_uint64)]),
    "gss_last_error": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libgss.so (built by ``make -C string-sorting_amd``)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise GssError(f"HIP library not built: {LIB_PATH} is missing "
                           "(run __graft_entry__.build() or make -C string-sorting_amd)")
        l = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def exported_symbols() -> list:
    return sorted(_SIGNATURES)


def last_error() -> str:
    return (lib().gss_last_error() or b"").decode()


def _check(rc: int, what: str) -> None:
    if rc != GSS_OK:
        raise GssError(f"{what} failed ({rc}): {last_error()}")


# ---------------------------------------------------------------------------
# Host pointer path (the routine.h drop-in)

def _pointer_array(bufs):
    arr = (C.c_void_p * max(1, len(bufs)))()
    for i, b in enumerate(bufs):
        arr[i] = C.addressof(b)
    return arr


def sort_pointers(ptrs, n: int, lcp: bool = False):
    """Calls gss_msd / gss_msd_lcp on a ctypes array of n char pointers."""
    if lcp:
        out = (C.c_size_t * max(1, n - 1))()
        lib().gss_msd_lcp(C.cast(ptrs, C.POINTER(C.c_void_p)), n, out)
        return list(out[: max(0, n - 1)])
    lib().gss_msd(C.cast(ptrs, C.POINTER(C.c_void_p)), n)
    return None


def set_pointer_limits(max_text_bytes: int = 0, max_strings: int = 0) -> None:
    """Largest device sort of the pointer entry points (0: the default,
    2^32-64 text bytes and 2^31-1 strings); larger inputs are split on the
    host by leading bytes first."""
    _check(lib().gss_set_pointer_limits(max_text_bytes, max_strings), "gss_set_pointer_limits")


def pointer_info() -> PointerInfo:
    """What the last gss_msd / gss_msd_lcp call did (bytes uploaded, device
    sorts, packed or address range copied)."""
    info = PointerInfo()
    _check(lib().gss_pointer_last_info(C.byref(info)), "gss_pointer_last_info")
    return info


def sort_strings(strs, lcp: bool = False):
    """Sorts byte strings through the pointer entry point; each string is a
    separate allocation, as in the reference's unit test (strdup).
    Returns the sorted strings (and the LCP list when lcp=True)."""
    bufs = [C.create_string_buffer(bytes(s)) for s in strs]
    ptrs = _pointer_array(bufs)
    n = len(bufs)
    l = sort_pointers(ptrs, n, lcp)
    addr = {C.addressof(b): i for i, b in enumerate(bufs)}
    order = [addr[ptrs[i]] for i in range(n)]
    out = [bytes(strs[i]) for i in order]
    return (out, l) if lcp else out


# ---------------------------------------------------------------------------
# Device-resident path

def _check_text(text, text_bytes):
    """The text must be a contiguous uint8 device tensor holding text_bytes
    bytes (default: all of it); returns text_bytes."""
    import torch

    if not text.is_cuda or text.dtype != torch.uint8 or not text.is_contiguous():
        raise GssError("text must be a contiguous uint8 device tensor")
    tb = text.numel() if text_bytes is None else int(text_bytes)
    if tb < 0 or tb > text.numel():
        raise GssError(f"text_bytes={tb} does not fit the {text.numel()}-byte text tensor")
    return tb


def _check_offsets(t, device, what, n=None):
    """A contiguous 32-bit tensor on `device` with at least n entries."""
    import torch

    if not t.is_cuda or t.device != device:
        raise GssError(f"{what} must be a device tensor on {device}")
    if t.dtype not in (torch.int32, torch.uint32) or not t.is_contiguous():
        raise GssError(f"{what} must be a contiguous 32-bit tensor")
    if n is not None and t.numel() < n:
        raise GssError(f"{what} must hold at least {n} entries (has {t.numel()})")


def _check_lcp(lcp, device, n):
    """A contiguous 64-bit tensor on `device` with at least n-1 entries."""
    if (not lcp.is_cuda or lcp.device != device or lcp.element_size() != 8
            or lcp.is_floating_point() or not lcp.is_contiguous()
            or lcp.numel() < max(0, n - 1)):
        raise GssError(f"lcp must be a contiguous 64-bit integer tensor of at least "
                       f"{max(0, n - 1)} entries on {device}")


def sort_device(text, offsets, out=None, lcp=None, text_bytes=None, stream=None):
    """Sorts the strings at uint32 `offsets` into uint8 `text`, both torch
    tensors on the GPU.  Writes the sorted offsets to `out` (allocated when
    None) and, when `lcp` is an int64/uint64 tensor of n-1 entries, the LCPs.
    Returns `out`."""
    import torch

    text_bytes = _check_text(text, text_bytes)
    _check_offsets(offsets, text.device, "offsets")
    n = offsets.numel()
    if out is None:
        out = torch.empty_like(offsets)
    _check_offsets(out, text.device, "out", n)
    if stream is None:
        stream = torch.cuda.current_stream(text.device)
    lp = 0
    if lcp is not None:
        _check_lcp(lcp, text.device, n)
        lp = lcp.data_ptr()
    rc = lib().gss_sort_device32(text.data_ptr(), int(text_bytes), offsets.data_ptr(),
                                 out.data_ptr(), lp if lp else None, n, stream.cuda_stream)
    _check(rc, "gss_sort_device32")
    return out


def _req_table(req_begin):
    arr = (C.c_uint64 * len(req_begin))(*[int(x) for x in req_begin])
    return arr


def sort_batch(text, offsets, req_begin, out=None, lcp=None, text_bytes=None, stream=None,
               ctx=None):
    """Sorts independent requests in one batch: request r owns offsets
    [req_begin[r], req_begin[r+1]) of the uint32 `offsets` tensor (all into
    `text`).  Each request's sorted offsets land at its own positions of
    `out`.  `ctx` is a Context (None = the default context)."""
    import torch

    text_bytes = _check_text(text, text_bytes)
    _check_offsets(offsets, text.device, "offsets")
    n = offsets.numel()
    nreq = len(req_begin) - 1
    if nreq < 0 or int(req_begin[0]) != 0 or int(req_begin[-1]) != n:
        raise GssError("req_begin must run from 0 to the number of offsets")
    if out is None:
        out = torch.empty_like(offsets)
    _check_offsets(out, text.device, "out", n)
    if lcp is not None:
        _check_lcp(lcp, text.device, n)
    lp = lcp.data_ptr() if lcp is not None else None
    table = _req_table(req_begin)
    if ctx is None:
        if stream is None:
            stream = torch.cuda.current_stream(text.device)
        rc = lib().gss_sort_batch_device32(text.data_ptr(), int(text_bytes), offsets.data_ptr(),
                                           out.data_ptr(), lp, table, nreq, stream.cuda_stream)
    else:
        st = stream.cuda_stream if stream is not None else None
        rc = lib().gss_ctx_sort_batch_device32(ctx.handle, text.data_ptr(), int(text_bytes),
                                               offsets.data_ptr(), out.data_ptr(), lp, table,
                                               nreq, st)
    _check(rc, "gss_sort_batch_device32")
    return out


class Context:
    """A sort context: its own HIP stream and device workspace, so sorts on
    different contexts can run at the same time (gss_ctx_* in gss.h)."""

    def __init__(self):
        self.handle = lib().gss_ctx_create()
        if not self.handle:
            raise GssError(f"gss_ctx_create failed: {last_error()}")

    @property
    def stream_handle(self):
        return lib().gss_ctx_stream(self.handle)

    def stats(self) -> "Stats":
        s = Stats()
        _check(lib().gss_ctx_get_stats(self.handle, C.byref(s)), "gss_ctx_get_stats")
        return s

    def close(self):
        if self.handle:
            lib().gss_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_ptr(t, stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream(t.device)
    return s.cuda_stream


def _ptr(t):
    return t.data_ptr() if t is not None else None


def ingest(text, delim=b"\n", text_bytes=None, out=None, stream=None):
    """Device-side create_strings_delim: splits the uint8 device tensor
    `text` in place at every `delim` byte (each becomes a NUL unless it is
    NUL) and returns the int32 offsets of the strings (one per delimiter).
    With `out` (an int32 tensor as large as the string count) the split is
    one call; without it a counting call sizes the result first."""
    import torch

    d = delim[0] if isinstance(delim, (bytes, bytearray)) else int(delim)
    tb = text.numel() if text_bytes is None else int(text_bytes)
    n = C.c_uint64()
    st = _stream_ptr(text, stream)
    if out is None:
        _check(lib().gss_ingest_device32(text.data_ptr(), tb, d, None, 0, C.byref(n), st),
               "gss_ingest_device32")
        out = torch.empty(max(1, n.value), dtype=torch.int32, device=text.device)
    _check(lib().gss_ingest_device32(text.data_ptr(), tb, d, out.data_ptr(), out.numel(),
                                     C.byref(n), st), "gss_ingest_device32")
    return out[: n.value]


def suffixes(text_bytes, device, stream=None):
    """Device-side create_suffixes: int32 offsets 0..text_bytes-1."""
    import torch

    off = torch.empty(max(1, int(text_bytes)), dtype=torch.int32, device=device)
    _check(lib().gss_suffixes_device32(int(text_bytes), off.data_ptr(), _stream_ptr(off, stream)),
           "gss_suffixes_device32")
    return off[: int(text_bytes)]


def write(text, offsets, text_bytes=None, out=None, stream=None):
    """Device-side write_result: the strings at `offsets`, in order, each
    followed by '\\n', as a uint8 device tensor.  Without `out` the first
    attempt assumes the strings do not overlap (output <= text size) and
    retries with the exact size when they do."""
    import torch

    tb = text.numel() if text_bytes is None else int(text_bytes)
    n = offsets.numel()
    size = C.c_uint64()
    st = _stream_ptr(text, stream)
    if out is None:
        out = torch.empty(max(16, tb + 16), dtype=torch.uint8, device=text.device)
    rc = lib().gss_write_device32(text.data_ptr(), tb, offsets.data_ptr(), n, out.data_ptr(),
                                  out.numel(), C.byref(size), st)
    if rc == -4 and size.value > out.numel():  # GSS_E_TOO_LARGE: grow once
        out = torch.empty(size.value, dtype=torch.uint8, device=text.device)
        rc = lib().gss_write_device32(text.data_ptr(), tb, offsets.data_ptr(), n, out.data_ptr(),
                                      out.numel(), C.byref(size), st)
    _check(rc, "gss_write_device32")
    return out[: size.value]


def merge_lcp(text, a, a_lcp, b, b_lcp, text_bytes=None, want_lcp=True, stream=None):
    """LCP-aware merge of two sorted runs (int32 offset tensors with their
    int64 LCP tensors of n-1 entries); returns (offsets, lcp or None)."""
    import torch

    tb = text.numel() if text_bytes is None else int(text_bytes)
    na, nb = a.numel(), b.numel()
    out = torch.empty(max(1, na + nb), dtype=torch.int32, device=text.device)
    ol = torch.empty(max(1, na + nb - 1), dtype=torch.int64, device=text.device) if want_lcp else None
    _check(lib().gss_merge_lcp_device32(text.data_ptr(), tb, _ptr(a) if na else None,
                                        _ptr(a_lcp) if na > 1 else None, na,
                                        _ptr(b) if nb else None, _ptr(b_lcp) if nb > 1 else None,
                                        nb, out.data_ptr(), _ptr(ol), _stream_ptr(text, stream)),
           "gss_merge_lcp_device32")
    return out[: na + nb], (ol[: max(0, na + nb - 1)] if want_lcp else None)


def merge_lcp_kway(text, runs, run_lcps, text_bytes=None, want_lcp=True, stream=None):
    """K-way LCP merge of sorted runs (lists of int32 offset tensors and
    their int64 LCP tensors of n-1 entries); equal strings keep run order.
    Returns (offsets, lcp or None)."""
    import torch

    if len(runs) != len(run_lcps):
        raise GssError("one LCP tensor per run")
    tb = text.numel() if text_bytes is None else int(text_bytes)
    k = len(runs)
    lens = [r.numel() for r in runs]
    total = sum(lens)
    out = torch.empty(max(1, total), dtype=torch.int32, device=text.device)
    ol = torch.empty(max(1, total - 1), dtype=torch.int64, device=text.device) if want_lcp else None
    offs = (C.c_void_p * max(1, k))(*[r.data_ptr() if n else None for r, n in zip(runs, lens)])
    lcps = (C.c_void_p * max(1, k))(*[l.data_ptr() if n > 1 else None
                                      for l, n in zip(run_lcps, lens)])
    nlen = (C.c_uint64 * max(1, k))(*lens)
    _check(lib().gss_merge_lcp_kway_device32(text.data_ptr(), tb, offs, lcps, nlen, k,
                                             out.data_ptr(), _ptr(ol), _stream_ptr(text, stream)),
           "gss_merge_lcp_kway_device32")
    return out[:total], (ol[: max(0, total - 1)] if want_lcp else None)


class Server:
    """The served request loop (gss_server_* in gss.h) over one device text.

    ``submit`` queues a sort of a uint32 offset tensor (a contiguous device
    tensor, e.g. a slice of a pool) into an output tensor of the same size.
    With ``ready`` (a recorded ``torch.cuda.Event``) the request's device work
    waits for the work that event marks, e.g. the kernel that wrote the
    offsets on another stream; without it that work must have finished
    before the call (e.g. ``torch.cuda.synchronize()``).  Slots take queued
    requests in order, fold as many as fit into one batch and complete them
    together.  ``wait`` returns the request's latency in ms.
    """

    def __init__(self, text, text_bytes=None, hbm_budget=8 << 30, slots=2,
                 max_batch_requests=4096):
        if not text.is_cuda:
            raise GssError("the server text must be a device tensor")
        self.text = text  # kept alive for the server's lifetime
        cfg = ServerConfig(int(hbm_budget), int(slots), int(max_batch_requests))
        tb = text.numel() if text_bytes is None else int(text_bytes)
        self.handle = lib().gss_server_create(text.data_ptr(), tb, C.byref(cfg))
        if not self.handle:
            raise GssError(f"gss_server_create failed: {last_error()}")
        self._keep = {}

    def submit(self, offsets, out, lcp=None, ready=None) -> int:
        n = offsets.numel()
        _check_offsets(offsets, self.text.device, "offsets")
        _check_offsets(out, self.text.device, "out", n)
        if lcp is not None:
            _check_lcp(lcp, self.text.device, n)
        ev = None
        if ready is not None:
            ev = getattr(ready, "cuda_event", 0) or None
            if ev is None:
                raise GssError("ready must be a recorded torch.cuda.Event")
        rid = lib().gss_server_submit(self.handle, offsets.data_ptr(), out.data_ptr(),
                                      lcp.data_ptr() if lcp is not None else None, n, ev)
        if rid < 0:
            raise GssError(f"gss_server_submit failed ({rid}): {last_error()}")
        self._keep[rid] = (offsets, out, lcp)
        return rid

    def wait(self, rid: int) -> float:
        ms = C.c_double()
        rc = lib().gss_server_wait(self.handle, rid, C.byref(ms))
        self._keep.pop(rid, None)
        _check(rc, "gss_server_wait")
        return ms.value

    def done(self, rid: int) -> bool:
        rc = lib().gss_server_done(self.handle, rid)
        if rc < 0:
            raise GssError(f"gss_server_done failed ({rc}): {last_error()}")
        return rc == 1

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib().gss_server_get_stats(self.handle, C.byref(s)), "gss_server_get_stats")
        return s

    def close(self):
        if self.handle:
            lib().gss_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(text, offsets, text_bytes=None, stream=None) -> AutotuneResult:
    """Times every candidate tuning on a sample of this input and returns
    the fastest (gss_autotune_device32); the current tuning is unchanged."""
    r = AutotuneResult()
    tb = text.numel() if text_bytes is None else int(text_bytes)
    _check(lib().gss_autotune_device32(text.data_ptr(), tb, offsets.data_ptr(), offsets.numel(),
                                       C.byref(r), _stream_ptr(text, stream)),
           "gss_autotune_device32")
    return r


def autotune_enable(on: bool = True) -> None:
    _check(lib().gss_autotune_enable(1 if on else 0), "gss_autotune_enable")


def autotune_reset() -> None:
    _check(lib().gss_autotune_reset(), "gss_autotune_reset")


def autotune_cached() -> int:
    return int(lib().gss_autotune_cached())


def get_tuning() -> Tuning:
    t = Tuning()
    _check(lib().gss_get_tuning(C.byref(t)), "gss_get_tuning")
    return t


def set_tuning(local_threshold=None, radix_bits=None, tile=None, key_bytes=None) -> Tuning:
    t = get_tuning()
    if local_threshold is not None:
        t.local_threshold = local_threshold
    if radix_bits is not None:
        t.radix_bits = radix_bits
    if tile is not None:
        t.tile = tile
    if key_bytes is not None:
        t.key_bytes = key_bytes
    _check(lib().gss_set_tuning(C.byref(t)), "gss_set_tuning")
    return t


def stats() -> Stats:
    s = Stats()
    _check(lib().gss_get_stats(C.byref(s)), "gss_get_stats")
    return s


def profile_enable(on: bool = True, levels: bool = False, passes_only: bool = False) -> None:
    """Kernel timing (HIP event brackets around every launch and pass) on or
    off; `passes_only` brackets only the passes and the local phase (an event
    between two kernels delays the second, so pass times are read this way);
    `levels` runs the local kernels' counting variant, which fills
    Stats.local_level_strings."""
    mode = (1 if on and not passes_only else 0) | (4 if on and passes_only else 0)
    _check(lib().gss_profile_enable(mode | (2 if levels else 0)), "gss_profile_enable")


def profile_reset() -> None:
    _check(lib().gss_profile_reset(), "gss_profile_reset")


def profile_get(name: str):
    """(launches, total_ms, items) accumulated for kernel `name`."""
    launches, ms, items = C.c_uint64(), C.c_double(), C.c_uint64()
    _check(lib().gss_profile_get(name.encode(), C.byref(launches), C.byref(ms), C.byref(items)),
           "gss_profile_get")
    return launches.value, ms.value, items.value


# ---------------------------------------------------------------------------
# Routine registry, mirroring src/routine.h:32-59 and src/routines.c:31-75.

@dataclass(frozen=True)
class Routine:
    name: str
    desc: str
    multicore: bool
    lcp: bool

    def f(self, strs):
        """Sorts `strs` (bytes) through this routine's C entry point."""
        out = sort_strings(strs, lcp=self.lcp)
        return out[0] if self.lcp else out


ROUTINES_MAX = 256
_routines = []


def routine_register(r: Routine) -> None:
    """routines.c:31-43: rejects missing fields and registry overflow."""
    if r is None or not r.name or not r.desc:
        raise GssError("routine_register: NULL routine, name or description")
    if len(_routines) >= ROUTINES_MAX:
        raise GssError("routine_register: registry full")
    _routines.append(r)


def routine_from_name(name: str):
    """routines.c:45-53: first routine with this name, or None."""
    for r in _routines:
        if r.name == name:
            return r
    return None


def routine_get_all():
    """routines.c:55-75: sorted by (multicore, name)."""
    return sorted(_routines, key=lambda r: (r.multicore, r.name))


routine_register(Routine("gss_msd", "MI355X MSD radix sort, 8-byte keys (HIP, gfx950)", True, False))
routine_register(Routine("gss_msd_with_lcp", "MI355X MSD radix sort + LCP array (HIP, gfx950)",
                         True, True))
