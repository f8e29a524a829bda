"""The C ABI library without a GPU: it loads, exports every symbol
include/gss.h declares, validates arguments, and does not depend on the
oracle.  CPU only: no call here reaches a device."""

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gss

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gss.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gss_\w+)\s*\(", text)))


def test_library_loads():
    assert gss.LIB_PATH.exists(), "build() must produce libgss.so"
    gss.lib()


def test_every_declared_symbol_is_exported():
    decl = declared_functions()
    assert decl, "no declarations parsed"
    lib = gss.lib()
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(gss.exported_symbols()) == decl
    out = subprocess.run(["nm", "-D", "--defined-only", str(gss.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gss_\w+)", out))
    assert set(decl) <= exported


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", str(gss.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    assert "oracle" not in out and "libref" not in out
    for src in (ROOT / "string-sorting_amd").rglob("*"):
        if src.suffix in (".hip", ".cpp", ".h", ".c", ".py") and "oracle" in src.read_text():
            body = src.read_text()
            assert "import oracle" not in body and "oracle.h" not in body, src


def test_tuning_validation():
    t = gss.get_tuning()
    try:
        with pytest.raises(gss.GssError):
            gss.set_tuning(local_threshold=10)
        with pytest.raises(gss.GssError):
            gss.set_tuning(local_threshold=20000)
        with pytest.raises(gss.GssError):
            gss.set_tuning(radix_bits=9)
        with pytest.raises(gss.GssError, match="tile"):
            gss.set_tuning(tile=8192)
        with pytest.raises(gss.GssError, match="key_bytes"):
            gss.set_tuning(key_bytes=16)
        gss.set_tuning(local_threshold=1024, radix_bits=11, tile=2048, key_bytes=32)
        assert (gss.get_tuning().tile, gss.get_tuning().key_bytes) == (2048, 32)
        assert gss.get_tuning().radix_bits == 11
    finally:
        gss.set_tuning(t.local_threshold, t.radix_bits, t.tile, t.key_bytes)


def test_pointer_limits_validation():
    with pytest.raises(gss.GssError, match="max_text_bytes"):
        gss.set_pointer_limits(max_text_bytes=63)
    with pytest.raises(gss.GssError, match="max_text_bytes"):
        gss.set_pointer_limits(max_text_bytes=1 << 32)
    with pytest.raises(gss.GssError, match="max_strings"):
        gss.set_pointer_limits(max_strings=1)
    with pytest.raises(gss.GssError, match="max_strings"):
        gss.set_pointer_limits(max_strings=1 << 31)
    gss.set_pointer_limits(4096, 100)
    gss.set_pointer_limits()  # back to the defaults
    info = gss.pointer_info()  # nothing sorted yet in this process: all zero
    assert (info.uploaded_text_bytes, info.pieces, info.packed) == (0, 0, 0)


def test_device_entry_rejects_host_tensors():
    torch = pytest.importorskip("torch")
    text = torch.zeros(16, dtype=torch.uint8)
    off = torch.zeros(1, dtype=torch.int32)
    with pytest.raises(gss.GssError):
        gss.sort_device(text, off)


def test_registry_mirrors_reference():
    """routines.c:31-75: lookup by name, listing sorted by (multicore, name),
    registration rejects incomplete routines."""
    assert gss.routine_from_name("gss_msd").multicore
    assert gss.routine_from_name("no_such_routine") is None
    names = [r.name for r in gss.routine_get_all()]
    assert names == sorted(names)
    with pytest.raises(gss.GssError):
        gss.routine_register(gss.Routine("", "x", False, False))
    with pytest.raises(gss.GssError):
        gss.routine_register(gss.Routine("x", "", False, False))


def test_stats_and_profile_without_device():
    s = gss.stats()
    assert s.global_passes >= 0
    gss.profile_reset()
    assert gss.profile_get("scatter") == (0, 0.0, 0)
    lib = gss.lib()
    assert lib.gss_profile_enable(8) == -1  # GSS_E_INVALID: unknown mode bit
    assert b"mode" in lib.gss_last_error()
    assert lib.gss_profile_enable(4) == 0  # GSS_PROFILE_PASSES
    gss.profile_enable(True, levels=True)
    gss.profile_enable(True, passes_only=True)
    gss.profile_enable(False)


def test_server_create_validates_before_touching_a_device():
    """Argument checks and HBM-budget admission of gss_server_create run on
    the host: a budget below one slot's workspace is refused with a message,
    as are slot counts outside [1, 8], before any device call."""
    import ctypes as C

    lib = gss.lib()
    fake_text = C.c_void_p(1 << 20)  # never dereferenced: every case fails first
    for budget, slots, maxreq, msg in [
        (1 << 30, 0, 16, "slots"),
        (1 << 30, 9, 16, "slots"),
        (1 << 30, 2, 0, "max_batch_requests"),
        (1000, 1, 16, "too small"),
    ]:
        cfg = gss.ServerConfig(budget, slots, maxreq)
        assert not lib.gss_server_create(fake_text, 4096, C.byref(cfg))
        assert msg in gss.last_error()
    cfg = gss.ServerConfig(1 << 30, 1, 16)
    assert not lib.gss_server_create(C.c_void_p(1 << 20 | 4), 4096, C.byref(cfg))
    assert "aligned" in gss.last_error()
    assert lib.gss_server_submit(None, None, None, None, 5, None) < 0
    assert lib.gss_server_wait(None, 0, None) < 0
    assert lib.gss_server_done(None, 0) < 0
    lib.gss_server_destroy(None)


def test_kway_merge_validates_before_touching_a_device():
    """The k-way merge checks its run table on the host: missing LCP arrays
    for runs of two or more strings and NULL runs are refused; no strings at
    all is a no-op."""
    import ctypes as C

    lib = gss.lib()
    lens = (C.c_uint64 * 2)(5, 3)
    offs = (C.c_void_p * 2)(1 << 20, 1 << 21)
    lcps = (C.c_void_p * 2)(1 << 22, None)
    assert lib.gss_merge_lcp_kway_device32(C.c_void_p(1 << 23), 64, offs, lcps, lens, 2,
                                           C.c_void_p(1 << 24), None, None) < 0
    assert "LCP array" in gss.last_error()
    offs2 = (C.c_void_p * 2)(None, 1 << 21)
    lcps2 = (C.c_void_p * 2)(1 << 22, 1 << 22)
    assert lib.gss_merge_lcp_kway_device32(C.c_void_p(1 << 23), 64, offs2, lcps2, lens, 2,
                                           C.c_void_p(1 << 24), None, None) < 0
    assert "NULL run" in gss.last_error()
    zero = (C.c_uint64 * 2)(0, 0)
    assert lib.gss_merge_lcp_kway_device32(None, 0, offs, lcps, zero, 2, None, None, None) == 0


def test_autotune_switches_without_device():
    """Enabling, resetting and counting the autotuner's cache touch no
    device; tuning rejects missing inputs before any device call."""
    import ctypes as C

    gss.autotune_reset()
    assert gss.autotune_cached() == 0
    gss.autotune_enable(True)
    gss.autotune_enable(False)
    r = gss.AutotuneResult()
    lib = gss.lib()
    assert lib.gss_autotune_device32(None, 0, None, 10, C.byref(r), None) < 0
    assert lib.gss_autotune_device32(C.c_void_p(1 << 20), 64, C.c_void_p(1 << 21), 1,
                                     C.byref(r), None) < 0
    assert "fewer than two" in gss.last_error()


def test_noalloc_merge_validates_before_touching_a_device():
    """The allocation-free merge checks its arguments and its size limit on
    the host: more strings than its fixed scratch covers is GSS_E_TOO_LARGE
    with a message naming the limit, an empty text or a missing LCP array is
    GSS_E_INVALID, and no strings at all is a no-op.  None of these reach a
    device (the pointers are never dereferenced)."""
    import ctypes as C

    lib = gss.lib()
    fake = C.c_void_p(1 << 20)
    na = nb = 300_000_000  # 6e8 strings > 536,868,864
    assert lib.gss_merge_lcp_device32_noalloc(fake, 4096, fake, fake, na, fake, fake, nb, fake,
                                              None, None) == -4
    assert "at most 536868864 strings" in gss.last_error()
    assert lib.gss_merge_lcp_device32_noalloc(fake, 0, fake, fake, 5, fake, fake, 5, fake,
                                              None, None) == -1
    assert "empty text" in gss.last_error()
    assert lib.gss_merge_lcp_device32_noalloc(fake, 4096, fake, None, 5, fake, fake, 5, fake,
                                              None, None) == -1
    assert "LCP" in gss.last_error()
    assert lib.gss_merge_lcp_device32_noalloc(None, 0, None, None, 0, None, None, 0, None,
                                              None, None) == 0
