"""Writes tests/golden/runtime_vectors.json: input/output vectors of the reference's
own CPU runtime helpers -- Hash, StrHash, StrCmp (com/cpp/cpu-runtime.hpp:37-78),
the fill + row-order insert of BuildHashTable / InsertHashTable
(com/cpp/com.hpp:168-174, com/cpp/cpu-runtime.hpp:80-103) and Probe<int> /
Probe<char> (com/cpp/cpu-runtime.hpp:125-152) -- computed by oracle/_ref/libref_rt.so,
which oracle/Makefile compiles from the reference's sources where they lie.

The vectors let the oracle be checked against the reference's runtime where those
sources (and so oracle/_ref) are absent.  Run from the repository root after
`make -C oracle`:

    python tests/golden/make_runtime_vectors.py
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_rt.so")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runtime_vectors.json")


def load():
    L = ctypes.CDLL(REF)
    for n in ("ref_hash", "ref_strhash"):
        f = getattr(L, n)
        f.restype, f.argtypes = ctypes.c_uint, [ctypes.c_char_p, ctypes.c_int]
    L.ref_strcmp.restype = ctypes.c_int
    L.ref_strcmp.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.ref_build_int.restype = ctypes.c_uint64
    L.ref_build_int.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint]
    L.ref_build_str.restype = ctypes.c_uint64
    L.ref_build_str.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint]
    L.ref_probe_int.restype = ctypes.c_int
    L.ref_probe_int.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    L.ref_probe_str.restype = ctypes.c_int
    L.ref_probe_str.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    return L


def main():
    L = load()
    rng = np.random.default_rng(20261204)
    out = {"source": "oracle/_ref/libref_rt.so built from com/cpp/cpu-runtime.hpp", "hash": [], "strcmp": [],
           "int_tables": [], "str_tables": []}
    # byte strings: random bytes (signed-char arithmetic), NULs inside, SSB-like text
    texts = [b"ASIA\0\0\0\0\0\0", b"UNITED KI1", b"MFGR#2221\0", b"NATION_07\0", b"\0abc", b"\xff\xfe\x80\x7f"]
    for _ in range(400):
        n = int(rng.integers(1, 16))
        b = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        if rng.random() < 0.2:
            b = b[: n // 2] + b"\0" + b[n // 2 + 1:]
        texts.append(b)
    for b in texts:
        out["hash"].append([b.hex(), L.ref_hash(b, len(b)), L.ref_strhash(b, len(b))])
    # StrCmp: equal, prefix (NUL in either side), differing bytes of either sign
    for _ in range(400):
        n = int(rng.integers(1, 12))
        a = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        c = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        r = rng.random()
        if r < 0.3:
            c = a
        elif r < 0.6:
            k = int(rng.integers(0, n))
            c = a[:k] + c[k:]
        if rng.random() < 0.25:
            k = int(rng.integers(0, n))
            if rng.random() < 0.5:
                a = a[:k] + b"\0" + a[k + 1:]
            else:
                c = c[:k] + b"\0" + c[k + 1:]
        out["strcmp"].append([a.hex(), c.hex(), n, L.ref_strcmp(a, c, n)])
    # int tables: sparse, crowded (drops), duplicates (first row wins), negative keys
    cases = [(1000, rng.permutation(np.arange(1, 3000, dtype=np.int32))[:300]),
             (1000, rng.permutation(np.arange(1, 3000, dtype=np.int32))[:990]),
             (64, rng.integers(-5, 20, 60).astype(np.int32)),
             (997, rng.integers(-(1 << 31), (1 << 31) - 1, 900, dtype=np.int64).astype(np.int32))]
    for slots, keys in cases:
        keys = np.ascontiguousarray(keys, np.int32)
        ht = np.empty(slots, np.int32)
        d = L.ref_build_int(keys.ctypes.data, len(keys), ht.ctypes.data, slots)
        probes = np.concatenate([keys[rng.integers(0, len(keys), 150)],
                                 rng.integers(-10, 4000, 100).astype(np.int32)]).astype(np.int32)
        ans = [L.ref_probe_int(int(k), keys.ctypes.data, ht.ctypes.data, slots) for k in probes.tolist()]
        out["int_tables"].append({"slots": slots, "keys": keys.tolist(), "dropped": int(d), "table": ht.tolist(),
                                  "probes": probes.tolist(), "found": ans})
    # string tables: fixed-width NUL-padded names (a crowded one drops inserts)
    for slots, n, width in [(1000, 250, 10), (300, 290, 10), (500, 200, 15)]:
        names = [(b"CITY%05d" % int(v)).ljust(width, b"\0")[:width] for v in rng.integers(0, 100000, n)]
        tab = np.frombuffer(b"".join(names), np.uint8).copy()
        ht = np.empty(slots, np.int32)
        d = L.ref_build_str(tab.ctypes.data, n, width, ht.ctypes.data, slots)
        probes = [names[int(i)] for i in rng.integers(0, n, 100)] + \
                 [(b"CITY%05d" % int(v)).ljust(width, b"\0")[:width] for v in rng.integers(0, 100000, 60)]
        ans = [L.ref_probe_str(p, width, tab.ctypes.data, ht.ctypes.data, slots) for p in probes]
        out["str_tables"].append({"slots": slots, "width": width, "keys": tab.tobytes().hex(), "dropped": int(d),
                                  "table": ht.tolist(), "probes": [p.hex() for p in probes], "found": ans})
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(out['hash'])} hashes, {len(out['strcmp'])} StrCmp, "
          f"{len(out['int_tables'])} int tables, {len(out['str_tables'])} string tables")


if __name__ == "__main__":
    main()
