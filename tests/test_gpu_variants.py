"""Every launch configuration the autotuners can choose answers like the oracle.

The plain scan's autotuner times candidate launch shapes of `svm_scan` (workgroup
size, unroll, maps in LDS or L2, prefetch, the flat form, workgroups per CU) and the
batched path's times every `k_shared_scan` variant x tile size x key-bit pre-test;
each keeps the fastest.  A box may pick any of them, so each one is forced in turn
(SVM_SCAN_VARIANT / SVM_BATCH_VARIANT) and its answer compared bit for bit with the
oracle's.  The tuners themselves are then run on these test-sized tables
(SVM_TUNE_MIN_ROWS / SVM_BATCH_TUNE_MIN_ROWS): every trial's group table is compared
on the device with the first trial's, and no trial may differ.
"""
import ctypes
import json

import pytest

from shadowvm_amd import _lib
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import (assert_same_columns, edge_query, host_tables, oracle_query,
                      oracle_slice_queries)

pytestmark = pytest.mark.gpu

# kTuneCandidates in shadowvm_amd/csrc/exec.cpp, kVariants in shadowvm_amd/csrc/batch_kernels.hip;
# tests/test_boundary.py checks both against the library (svm_debug_variant_counts)
N_SCAN_CANDIDATES = 50
N_BATCH_VARIANTS = 16


def tune_report(reset=False):
    L = ctypes.CDLL(_lib.LIB_PATH)
    f = L.svm_debug_tune_report
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int]
    buf = ctypes.create_string_buffer(1 << 20)
    rc = f(buf, len(buf), 1 if reset else 0)
    assert rc == 0, L.svm_last_error
    return json.loads(buf.value.decode())


def _same(got, want):
    assert got.kind == want.kind
    assert got.rows == want.rows, (got.rows, want.rows)
    assert_same_columns(got.columns, want.columns)


def _plain_workloads():
    return [
        ("four_way", ssb.four_way(seed=13, fact_rows=4_000_001, n_cust=300_000, n_supp=20_000, n_part=140_000,
                                  prefix="vf")),
        ("three_way", ssb.three_way(seed=17, fact_rows=4_000_003, n_part=100_000, n_supp=10_000, prefix="vt")),
        ("single_join", ssb.single_join(seed=19, fact_rows=4_000_002, prefix="vs")),
    ]


@pytest.mark.parametrize("name", ["four_way", "three_way", "single_join"])
def test_every_scan_variant_matches_oracle(engine, monkeypatch, name):
    wl = dict(_plain_workloads())[name]
    print(f"[variants] {name}: oracle on {wl.fact_rows} rows", flush=True)
    want, _ = oracle_query(wl, host_tables(wl))
    ssb.load_device(wl)
    ssb.run_dims(wl)
    for v in range(N_SCAN_CANDIDATES):
        monkeypatch.setenv("SVM_SCAN_VARIANT", str(v))
        print(f"[variants] {name}: svm_scan candidate {v}", flush=True)
        _same(ssb.run_fact(wl), want)
    monkeypatch.delenv("SVM_SCAN_VARIANT")
    with pytest.raises(_lib.SvmError):
        monkeypatch.setenv("SVM_SCAN_VARIANT", str(N_SCAN_CANDIDATES))
        ssb.run_fact(wl)
    monkeypatch.delenv("SVM_SCAN_VARIANT")
    # the tuner's own trials on this shape: none may differ from the first
    tune_report(reset=True)
    monkeypatch.setenv("SVM_TUNE_MIN_ROWS", "1")
    _same(ssb.run_fact(wl), want)
    rep = [r for r in tune_report()["scan"] if r["pipeline"] == wl.fact[0].pipe_info["id"]]
    print(f"[variants] {name}: tuner chose candidate {rep[0]['best'] if rep else None}", flush=True)
    assert len(rep) == 1 and len(rep[0]["trials"]) == N_SCAN_CANDIDATES
    assert rep[0]["rejected"] == 0, rep[0]
    # the tuned choice is cached and answers the same
    _same(ssb.run_fact(wl), want)
    for t in wl.tables:
        P.table_drop(t.id)


# kTuneCandidates' last ten stream the tier-0 key column from its narrow copy
PACKED_CANDIDATES = range(N_SCAN_CANDIDATES - 10, N_SCAN_CANDIDATES)


def scan_config(pipeline_id):
    L = ctypes.CDLL(_lib.LIB_PATH)
    f = L.svm_debug_scan_config
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    buf = ctypes.create_string_buffer(4096)
    assert f(pipeline_id.encode(), buf, len(buf)) == 0
    return buf.value.decode()


def test_packed_candidates_at_every_copy_width(engine, monkeypatch):
    """The narrow copy of the tier-0 key column holds 16, 18 or 24 bits a row (the
    fewest its range needs, or at least SVM_PACK_HIGH_BITS high bits): every packed
    candidate is checked against the oracle at each width.  Both filtering keys' ranges
    fit 16 bits, so the forced widths are the ones used."""
    wl = ssb.four_way(seed=29, fact_rows=4_000_005, n_cust=60_000, n_supp=20_000, n_part=140_000, prefix="vn")
    print(f"[variants] narrow copies: oracle on {wl.fact_rows} rows", flush=True)
    want, _ = oracle_query(wl, host_tables(wl))
    ssb.load_device(wl)
    ssb.run_dims(wl)
    pid = wl.fact[0].pipe_info["id"]
    for high, bits in (("0", 16), ("2", 18), ("8", 24)):
        monkeypatch.setenv("SVM_PACK_HIGH_BITS", high)
        for v in PACKED_CANDIDATES:
            monkeypatch.setenv("SVM_SCAN_VARIANT", str(v))
            print(f"[variants] narrow copies: {bits} bits, svm_scan candidate {v}", flush=True)
            _same(ssb.run_fact(wl), want)
            cfg = scan_config(pid)
            assert f"packed{bits}(" in cfg, cfg
    for t in wl.tables:
        P.table_drop(t.id)


@pytest.fixture(scope="module")
def batch_case(engine):
    wl = ssb.four_way(seed=23, fact_rows=6_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000, prefix="vb")
    ssb.load_device(wl)
    n = wl.fact_rows
    # template queries plus a few with hundreds to thousands of groups (the LDS hash
    # tables fill and overflow into the HBM tables)
    qs = [
        edge_query(wl, "vbe0", 0, n, [("customer", 'StrCmp($(0.2),"NATION_03",10)!=0', "$(0.0), $(0.3)", 1)],
                   "", "$(1.1)", "$(0.4)+$(0.5)"),
        edge_query(wl, "vbe1", 3, n - 7,
                   [("customer", 'StrCmp($(0.3),"REGION_1",8)==0', "$(0.0), $(0.1)", 1),
                    ("supplier", "", "$(0.0), $(0.1)", 2)],
                   "$(0.5)>=1000", "$(1.1), $(2.1)", "$(0.5)-$(0.4)"),
        edge_query(wl, "vbe2", 1_000_003, 2_000_001,
                   [("date", "$(0.1)>=1993 && $(0.1)<=1995", "$(0.0), $(0.1)", 0),
                    ("part", "$(0.1)<=12", "$(0.0), $(0.2)", 3)],
                   "$(0.4)<=5000000", "$(1.1), $(2.1)", "$(0.4)*$(0.5)"),
    ]
    qs += ssb.batch_queries(wl, n=9, seed=29, window=n, min_rows=500_000, prefix="vbq")
    ssb.make_views(qs)
    tabs = host_tables(wl)
    print(f"[variants] batch: oracle for {len(qs)} queries", flush=True)
    want = oracle_slice_queries(qs, tabs, tabs[wl.fact[1]])
    print(f"[variants] batch: groups per query {[w.rows for w in want]}", flush=True)
    plans, pos = [], []
    for q in qs:
        plans += q.plans()
        pos.append(len(plans) - 1)
    return qs, want, plans, pos


def _check_batch(batch_case, tag):
    qs, want, plans, pos = batch_case
    res, fused = P.batch_run(plans)
    assert fused >= len(qs) - 2, (tag, fused)
    for q, w, k in zip(qs, want, pos):
        got = res[k]
        assert got is not None, (tag, q.qid)
        assert got.rows == w.rows, (tag, q.qid, got.rows, w.rows)
        assert_same_columns(got.columns, w.columns)


@pytest.mark.parametrize("variant", range(N_BATCH_VARIANTS))
def test_every_shared_scan_variant_matches_oracle(batch_case, monkeypatch, variant):
    for tile in (16384, 65536):
        for bits in (0, 1):
            monkeypatch.setenv("SVM_BATCH_VARIANT", f"{variant},{tile},{bits}")
            print(f"[variants] batch: variant {variant} tile {tile} key bits {bits}", flush=True)
            _check_batch(batch_case, (variant, tile, bits))


STAGED_BATCH_VARIANTS = (12, 13, 14, 15)


@pytest.mark.parametrize("variant", STAGED_BATCH_VARIANTS)
def test_staged_shared_scan_over_any_tier0(batch_case, monkeypatch, variant):
    """Staged loads with tier 0 forced to the first k slots' key columns (k = 0: no
    staging; k past the slot count: every slot in tier 0)."""
    for k in (0, 1, 2, 3, 6):
        monkeypatch.setenv("SVM_BATCH_VARIANT", f"{variant},16384,{k % 2}")
        monkeypatch.setenv("SVM_BATCH_TIER0_SLOTS", str(k))
        print(f"[variants] batch: staged variant {variant} tier-0 slots {k}", flush=True)
        _check_batch(batch_case, (variant, "tier0", k))


def test_batch_tuner_trials_agree(batch_case, monkeypatch):
    tune_report(reset=True)
    monkeypatch.setenv("SVM_BATCH_TUNE_MIN_ROWS", "1")
    _check_batch(batch_case, "tuned")
    rep = tune_report()["batch"]
    print(f"[variants] batch tuner: {[(r['variant'], r['tile'], r['bits']) for r in rep]}", flush=True)
    assert rep and all(r["trials"] for r in rep)
    assert all(r["rejected"] == 0 for r in rep), rep
    _check_batch(batch_case, "tuned-cached")
