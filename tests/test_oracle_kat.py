"""Pins the oracle's runtime primitives (Hash, StrHash, StrCmp, the 10-probe
insert and Probe) against closed forms, the known-answer values of the
reference's hash, published drop counts, and -- when oracle/_ref was built from
the reference's own com/cpp/cpu-runtime.hpp -- the reference functions
themselves on random inputs."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RT = os.path.join(ROOT, "oracle", "_build", "liboracle_rt.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_rt.so")


def _load(path, prefix):
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    for n in ("hash", "strhash"):
        f = getattr(L, f"{prefix}_{n}")
        f.restype, f.argtypes = ctypes.c_uint, [ctypes.c_char_p, ctypes.c_int]
    f = getattr(L, f"{prefix}_strcmp")
    f.restype, f.argtypes = ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    f = getattr(L, f"{prefix}_build_int")
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint]
    f = getattr(L, f"{prefix}_build_str")
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint]
    f = getattr(L, f"{prefix}_probe_int")
    f.restype, f.argtypes = ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    f = getattr(L, f"{prefix}_probe_str")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    return L


@pytest.fixture(scope="module")
def rt():
    if not os.path.exists(RT):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_build/liboracle_rt.so"], check=True)
    return _load(RT, "ort")


def i32(v):
    return int(v).to_bytes(4, "little", signed=True)


def pad(s, n):
    b = s.encode()
    return b + b"\0" * (n - len(b))


# Closed forms: a little-endian int32 k < 256 is bytes [k,0,0,0], so
# Hash = k*31^3; 256 is [0,1,0,0] -> 31^2; -1 is four 0xFF bytes read as -1.
def test_hash_closed_forms(rt):
    assert rt.ort_hash(i32(1), 4) == 31 ** 3
    assert rt.ort_hash(i32(7), 4) == 7 * 31 ** 3
    assert rt.ort_hash(i32(256), 4) == 31 ** 2
    assert rt.ort_hash(i32(65536), 4) == 31
    assert rt.ort_hash(i32(1 << 24), 4) == 1
    minus1 = (-(31 ** 3 + 31 ** 2 + 31 + 1)) % (1 << 32) & 0x7FFFFFFF
    assert rt.ort_hash(i32(-1), 4) == minus1
    # signed bytes: 255 is [-1,0,0,0]
    assert rt.ort_hash(i32(255), 4) == ((-1 * 31 ** 3) % (1 << 32)) & 0x7FFFFFFF


# Known answers of the reference hash (little-endian int32 keys, & 0x7FFFFFFF).
KAT_INT = [(1, 29791), (255, 2147453857), (256, 961), (19920101, 2146669217),
           (19981231, 2145044197), (1400000, 2145666087), (3000000, 2145522681), (-1, 2147452864)]


@pytest.mark.parametrize("key,h", KAT_INT)
def test_hash_kat_int(rt, key, h):
    assert rt.ort_hash(i32(key), 4) == h
    assert h % 10_000_000 == rt.ort_hash(i32(key), 4) % 10_000_000


def test_hash_kat_strings(rt):
    asia = pad("ASIA", 10)
    assert rt.ort_hash(asia, 10) == 645958986
    assert rt.ort_strhash(asia, 10) == 2018506
    uk = b"UNITED KI1"
    assert rt.ort_hash(uk, 10) == rt.ort_strhash(uk, 10) == 1032449558
    mf = pad("MFGR#2221", 10)
    assert rt.ort_hash(mf, 10) == 76643938
    assert rt.ort_strhash(mf, 10) == 972303710


def test_strcmp_prefix_semantics(rt):
    # equal up to the first NUL in either operand (com/cpp/cpu-runtime.hpp:37-45)
    assert rt.ort_strcmp(b"MFGR#12\0", b"MFGR#1\0\0", 8) == 0
    assert rt.ort_strcmp(b"", b"XYZ", 5) == 0
    assert rt.ort_strcmp(b"AB", b"AC", 2) == ord("B") - ord("C")
    assert rt.ort_strcmp(b"NATION_07\0", b"NATION_07", 10) == 0
    assert rt.ort_strcmp(b"NATION_07\0", b"NATION_08", 10) == -1
    # full-width values without NUL behave like memcmp's sign
    assert rt.ort_strcmp(b"UNITED KI1", b"UNITED KI2", 10) < 0


def _drops(rt, keys, slots):
    keys = np.ascontiguousarray(keys, np.int32)
    ht = np.empty(slots, np.int32)
    d = rt.ort_build_int(keys.ctypes.data, len(keys), ht.ctypes.data, slots)
    return int(d), ht


# Dropped inserts of primary keys 1..N inserted in order (the 10-probe cap).
@pytest.mark.parametrize("n,slots,want", [(2556, 10_000_000, 0), (1_400_000, 10_000_000, 6910),
                                          (1_400_000, 15_000_000, 0), (800_000, 10_000_000, 630),
                                          (800_000, 15_000_000, 0), (3_000_000, 15_000_000, 746_032)])
def test_drop_counts(rt, n, slots, want):
    d, _ = _drops(rt, np.arange(1, n + 1), slots)
    assert d == want


def test_probe_finds_inserted_and_misses_dropped(rt):
    n, slots = 800_000, 10_000_000
    keys = np.arange(1, n + 1, dtype=np.int32)
    d, ht = _drops(rt, keys, slots)
    inserted = set(ht[ht >= 0].tolist())
    assert len(inserted) == n - d
    rng = np.random.default_rng(0)
    for r in rng.choice(n, 2000, replace=False).tolist():
        got = rt.ort_probe_int(int(keys[r]), keys.ctypes.data, ht.ctypes.data, slots)
        assert got == (r if r in inserted else -1)
    assert rt.ort_probe_int(n + 5, keys.ctypes.data, ht.ctypes.data, slots) == -1


def test_probe_duplicate_keys_first_row_wins(rt):
    keys = np.array([5, 9, 5, 5, 9], np.int32)
    d, ht = _drops(rt, keys, 50)
    assert d == 0
    assert rt.ort_probe_int(5, keys.ctypes.data, ht.ctypes.data, 50) == 0
    assert rt.ort_probe_int(9, keys.ctypes.data, ht.ctypes.data, 50) == 1


def test_string_table_probe(rt):
    names = [pad(f"NATION_{k:02d}", 10) for k in range(25)]
    tab = np.frombuffer(b"".join(names), np.uint8).copy()
    ht = np.empty(1000, np.int32)
    assert rt.ort_build_str(tab.ctypes.data, 25, 10, ht.ctypes.data, 1000) == 0
    for k, nm in enumerate(names):
        assert rt.ort_probe_str(nm, 10, tab.ctypes.data, ht.ctypes.data, 1000) == k
    assert rt.ort_probe_str(pad("NOPE", 10), 10, tab.ctypes.data, ht.ctypes.data, 1000) == -1


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")
def test_against_reference_runtime(rt):
    ref = _load(REF, "ref")
    rng = np.random.default_rng(1)
    for _ in range(3000):
        n = int(rng.integers(1, 16))
        b = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        assert rt.ort_hash(b, n) == ref.ref_hash(b, n)
        assert rt.ort_strhash(b, n) == ref.ref_strhash(b, n)
        c = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        if rng.random() < 0.5:
            c = b[: n // 2] + c[n // 2:]
        assert np.sign(rt.ort_strcmp(b, c, n)) == np.sign(ref.ref_strcmp(b, c, n))
        assert rt.ort_strcmp(b, c, n) == ref.ref_strcmp(b, c, n)
    # tables: random subsets in row order, small and crowded
    for slots, n in [(1000, 900), (50_000, 40_000), (10_000_000, 1_400_000)]:
        keys = rng.permutation(np.arange(1, 3 * n, dtype=np.int32))[:n].copy()
        h1, h2 = np.empty(slots, np.int32), np.empty(slots, np.int32)
        d1 = rt.ort_build_int(keys.ctypes.data, n, h1.ctypes.data, slots)
        d2 = ref.ref_build_int(keys.ctypes.data, n, h2.ctypes.data, slots)
        assert d1 == d2 and np.array_equal(h1, h2)
        for k in rng.integers(0, 3 * n, 500).tolist():
            assert rt.ort_probe_int(k, keys.ctypes.data, h1.ctypes.data, slots) == \
                ref.ref_probe_int(k, keys.ctypes.data, h2.ctypes.data, slots)
