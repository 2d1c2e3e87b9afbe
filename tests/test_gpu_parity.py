"""GPU parity: libsvm results through the C ABI against the CPU oracle, bit-exact.

Every comparison covers the full output: group keys (int32) and int64 sums in
key order, the scalar row, and the compacted dimension objects byte for byte.
"""
import numpy as np
import pytest

from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query

pytestmark = pytest.mark.gpu


def _run_both(engine, wl, tuple_per_bulk=3_000_000):
    print(f"[parity] {wl.name} fact_rows={wl.fact_rows} constants={wl.constants}", flush=True)
    tabs = host_tables(wl, tuple_per_bulk=tuple_per_bulk)
    want, want_dims = oracle_query(wl, tabs, tuple_per_bulk=tuple_per_bulk)
    ssb.load_device(wl)
    for pt, src, refs in wl.dims:
        got = pt.Run(src, refs)
        w = want_dims[pt.pipe_info["id"]]
        assert got.rows == w.rows
        assert_same_columns(got.columns, w.columns)
    got = ssb.run_fact(wl)
    assert got.kind == want.kind
    assert got.rows == want.rows
    assert_same_columns(got.columns, want.columns)
    return got, want


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_single_join_scalar(engine, seed):
    wl = ssb.single_join(seed=seed, fact_rows=1_000_003, prefix=f"sj{seed}")
    got, want = _run_both(engine, wl)
    assert got.rows == 1 and got.columns[0][0] > 0


def test_three_way_group(engine):
    wl = ssb.three_way(seed=5, fact_rows=2_000_001, n_part=80_000, n_supp=2_000, prefix="tw5")
    got, _ = _run_both(engine, wl)
    assert 100 < got.rows <= 7 * 40


@pytest.mark.parametrize("flat", ["0", "1"])
def test_four_way_zipf_strings(engine, monkeypatch, flat):
    # both code forms of the scan: joins under per-row branches, and the flat form
    # (the four rows' filters and map lookups straight-line, then aggregation)
    monkeypatch.setenv("SVM_SCAN_FLAT", flat)
    wl = ssb.four_way(seed=7, fact_rows=3_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000,
                      prefix="fw7")
    got, _ = _run_both(engine, wl)
    assert got.rows > 0


@pytest.mark.parametrize("flat", ["0", "1"])
def test_three_way_forms(engine, monkeypatch, flat):
    # the three-way group-by in both code forms
    monkeypatch.setenv("SVM_SCAN_FLAT", flat)
    wl = ssb.three_way(seed=11, fact_rows=2_000_000, n_part=100_000, n_supp=10_000, prefix=f"tf{flat}")
    got, _ = _run_both(engine, wl)
    assert got.rows > 0


def test_int32_metric_wraps(engine):
    """Per-row int32 products that overflow wrap on the GPU exactly as in the oracle
    (both compile with -fwrapv); the sums are int64."""
    from oracle import oracle as O

    from .test_oracle_brute import wrapping_plan, wrapping_table
    k, a, b = wrapping_table(n=2_000_003, seed=12)
    P.table_upload("wrap_src", [P.INT32] * 3, [4] * 3, [k, a, b])
    plan = wrapping_plan()
    want = O.run_pipeline(plan.plan("wrap_src", []), O.OTable([2, 2, 2], [4, 4, 4], [k, a, b]), [])
    got = plan.Run("wrap_src", [])
    assert got.rows == want.rows == 7
    assert_same_columns(got.columns, want.columns)
    P.table_drop("wrap_src")
