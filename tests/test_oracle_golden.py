"""The oracle's runtime primitives against golden vectors of the reference's own CPU
runtime (tests/golden/runtime_vectors.json, written by
tests/golden/make_runtime_vectors.py from oracle/_ref, which is compiled from
com/cpp/cpu-runtime.hpp): Hash / StrHash / StrCmp (cpu-runtime.hpp:37-78), the
fill + row-order insert of BuildHashTable (com/cpp/com.hpp:168-174) with its dropped
inserts, and Probe<int> / Probe<char> (cpu-runtime.hpp:125-152).  These hold where
the reference sources, and so oracle/_ref, are absent."""
import json
import os

import numpy as np
import pytest

from .test_oracle_kat import rt  # noqa: F401  (module fixture: the oracle's runtime library)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "runtime_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_hash_and_strhash(rt, golden):  # noqa: F811
    assert len(golden["hash"]) > 400
    for hx, h, sh in golden["hash"]:
        b = bytes.fromhex(hx)
        assert rt.ort_hash(b, len(b)) == h, hx
        assert rt.ort_strhash(b, len(b)) == sh, hx


def test_strcmp(rt, golden):  # noqa: F811
    nonzero = 0
    for ax, cx, n, want in golden["strcmp"]:
        assert rt.ort_strcmp(bytes.fromhex(ax), bytes.fromhex(cx), n) == want, (ax, cx, n)
        nonzero += want != 0
    assert 0 < nonzero < len(golden["strcmp"])  # both outcomes are covered


def test_int_tables_and_probes(rt, golden):  # noqa: F811
    dropped_any = False
    for t in golden["int_tables"]:
        keys = np.array(t["keys"], np.int32)
        ht = np.empty(t["slots"], np.int32)
        d = rt.ort_build_int(keys.ctypes.data, len(keys), ht.ctypes.data, t["slots"])
        assert d == t["dropped"]
        assert ht.tolist() == t["table"]
        dropped_any |= d > 0
        for k, want in zip(t["probes"], t["found"]):
            assert rt.ort_probe_int(k, keys.ctypes.data, ht.ctypes.data, t["slots"]) == want, k
    assert dropped_any


def test_string_tables_and_probes(rt, golden):  # noqa: F811
    for t in golden["str_tables"]:
        tab = np.frombuffer(bytes.fromhex(t["keys"]), np.uint8).copy()
        n = len(tab) // t["width"]
        ht = np.empty(t["slots"], np.int32)
        assert rt.ort_build_str(tab.ctypes.data, n, t["width"], ht.ctypes.data, t["slots"]) == t["dropped"]
        assert ht.tolist() == t["table"]
        for px, want in zip(t["probes"], t["found"]):
            assert rt.ort_probe_str(bytes.fromhex(px), t["width"], tab.ctypes.data, ht.ctypes.data,
                                    t["slots"]) == want, px
