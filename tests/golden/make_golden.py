"""Writes tests/golden/golden.json from the reference's own compiled routines.

Run in the container that has the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden.py

For every workload at a few small sizes it records the generator digest, and
the digests of the sorted output of the reference's msd_CE2 (offset order,
bytes, LCP array) and of multikey_cache8 (bytes).  For tiny sets it stores the
full input and output.  The reference publishes no golden files of its own
(its unit test builds its inputs inline, unit-test/main.cpp:207-272), so these
vectors are produced here from its code.  The LCP digests use the closed form
of lcp() (src/mergesort_lcp.cpp:88-100) evaluated by brute force in Python on
the reference's output order.
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "string-sorting_amd"), str(ROOT / "oracle")]

import gss.gen as gen  # noqa: E402
import oracle  # noqa: E402

SIZES = [2000, 50000]
TINY = 48


def strings_at(ss, off):
    t = ss.text.tobytes()
    return [t[o:t.index(b"\0", o)] for o in off.tolist()]


def lcp_brute(strs):
    out = []
    for a, b in zip(strs[:-1], strs[1:]):
        i = 0
        while i < len(a) and i < len(b) and a[i] == b[i]:
            i += 1
        out.append(i)
    return out


def digest_bytes(strs):
    return hashlib.md5(b"\n".join(strs) + b"\n").hexdigest()


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: run make -C oracle ref first")
    out = {"workloads": [], "tiny": []}
    for config in range(4):
        for n in SIZES:
            ss = gen.make(config, n=n)
            ce2 = oracle.ref_sort_offsets(ss.text, ss.offsets, "msd_CE2")
            mk = oracle.ref_sort_offsets(ss.text, ss.offsets, "multikey_cache8")
            s_ce2 = strings_at(ss, ce2)
            s_mk = strings_at(ss, mk)
            lcp = np.array(lcp_brute(s_ce2), dtype="<u8")
            out["workloads"].append({
                "config": config, "name": gen.NAMES[config], "n": n, "seed": gen.SEEDS[config],
                "input_md5": ss.md5(),
                "msd_CE2_offsets_md5": hashlib.md5(ce2.astype("<u4").tobytes()).hexdigest(),
                "msd_CE2_bytes_md5": digest_bytes(s_ce2),
                "multikey_cache8_bytes_md5": digest_bytes(s_mk),
                "lcp_md5": hashlib.md5(lcp.tobytes()).hexdigest(),
                "lcp_sum": int(lcp.sum()),
            })
            print(config, n, out["workloads"][-1]["msd_CE2_bytes_md5"])
        ss = gen.make(config, n=TINY, seed=gen.SEEDS[config] + 1)
        ce2 = oracle.ref_sort_offsets(ss.text, ss.offsets, "msd_CE2")
        s_in = ss.strings()
        s_out = strings_at(ss, ce2)
        index = {int(o): i for i, o in enumerate(ss.offsets.tolist())}
        out["tiny"].append({
            "config": config, "seed": gen.SEEDS[config] + 1,
            "input": [s.decode("latin-1") for s in s_in],
            "msd_CE2_order": [index[int(o)] for o in ce2.tolist()],
            "lcp": lcp_brute(s_out),
        })
    path = Path(__file__).with_name("golden.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
