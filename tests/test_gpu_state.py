"""State of the data plane across calls, on the GPU, against the oracle:

* a column rewritten with svm_table_put_column after queries have read it: facts
  cached about the old bytes (key ranges, dictionaries, kept rows, hash tables) are
  recomputed, and a dimension object materialised from the old bytes keeps
  answering from its own rows;
* served queries that carry the same pipeline ids (as the reference's drivers do,
  test/test-SSB-Q21.cu:21) run concurrently without touching each other's
  intermediate objects, in both serving modes;
* stopping the server while another thread is blocked waiting on a ticket, and
  waiting on tickets after the stop.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import _lib
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query
from .test_oracle_brute import wrapping_plan, wrapping_table

pytestmark = pytest.mark.gpu


def _same(got, want):
    assert got.rows == want.rows, (got.rows, want.rows)
    assert_same_columns(got.columns, want.columns)


def _rewrite(tabs, tid, col, values):
    """Column `col` of table `tid` rewritten on the device and in the oracle's copy."""
    t = tabs[tid]
    cols = list(t.columns)
    cols[col] = np.ascontiguousarray(values)
    tabs[tid] = O.OTable(t.types, t.sizes, cols, max_tuple_count=t.max_tuple_count)
    P.table_put(tid, col, cols[col])


def test_put_column_after_queries(engine):
    wl = ssb.single_join(seed=4, fact_rows=1_000_001, prefix="pc")
    date_id, fact_id = wl.tables[0].id, wl.tables[1].id
    tabs = host_tables(wl)
    ssb.load_device(wl)
    want, _ = oracle_query(wl, tabs)
    _same(ssb.run_query(wl), want)
    # every date key moves up by one: the key range and the key -> row mapping change
    _rewrite(tabs, date_id, 0, tabs[date_id].columns[0] + 1)
    # the dimension object made from the old bytes still holds the old keys
    _same(ssb.run_fact(wl), want)
    # dimension pipeline and query again over the new bytes
    want2, _ = oracle_query(wl, tabs)
    assert want2.columns[0][0] != want.columns[0][0]
    _same(ssb.run_query(wl), want2)
    # a fact column rewritten too (discount: the filter's column and a metric operand)
    _rewrite(tabs, fact_id, 2, (tabs[fact_id].columns[2] + 3) % 11)
    want3, _ = oracle_query(wl, tabs)
    assert want3.columns[0][0] != want2.columns[0][0]
    _same(ssb.run_query(wl), want3)
    for t in wl.tables:
        P.table_drop(t.id)


def test_put_group_key_column_after_queries(engine):
    """A source-column group key's range is cached per table; rewriting the column
    with values outside the old range gives the new groups."""
    k, a, b = wrapping_table(n=300_000, seed=9)
    P.table_upload("pg_src", [P.INT32] * 3, [4] * 3, [k, a, b])
    plan = wrapping_plan()
    t = O.OTable([2, 2, 2], [4, 4, 4], [k, a, b], max_tuple_count=3_000_000)
    _same(plan.Run("pg_src", []), O.run_pipeline(plan.plan("pg_src", []), t, []))
    k2 = (k * 7 + 100).astype(np.int32)
    P.table_put("pg_src", 0, k2)
    t2 = O.OTable([2, 2, 2], [4, 4, 4], [k2, a, b], max_tuple_count=3_000_000)
    _same(plan.Run("pg_src", []), O.run_pipeline(plan.plan("pg_src", []), t2, []))
    P.table_drop("pg_src")


@pytest.fixture(scope="module")
def served(engine):
    wl = ssb.four_way(seed=51, fact_rows=3_000_000, n_cust=200_000, n_supp=20_000, n_part=100_000,
                      prefix="sid")
    want, _ = oracle_query(wl, host_tables(wl))
    ssb.load_device(wl)
    return wl, want


@pytest.mark.parametrize("mode", [1, 0])
def test_same_ids_served_concurrently(served, mode):
    wl, want = served
    plans = wl.plans()  # fixed ids: every submission names the same dimension objects
    P.server_start(workers=3, budget=0, max_batch=64, mode=mode)
    try:
        tickets = [P.server_submit(plans) for _ in range(9)]
        results = [P.server_wait(t, timeout_ms=120_000) for t in tickets]
    finally:
        P.server_stop()
    for r in results:
        _same(r, want)
    # the served queries left no objects under their ids
    for pt, _, _ in wl.dims:
        assert not P.table_exists(pt.pipe_info["id"])


def test_stop_while_a_wait_is_blocked(served):
    wl, want = served
    plans = wl.plans()
    P.server_start(workers=1, budget=0, max_batch=1, mode=1)
    stopped = False
    try:
        tickets = [P.server_submit(plans) for _ in range(12)]
        outcome = {}

        def waiter():
            try:
                outcome["result"] = P.server_wait(tickets[-1], timeout_ms=-1)
            except _lib.SvmError as e:
                outcome["code"] = e.code

        th = threading.Thread(target=waiter)
        th.start()
        first = P.server_wait(tickets[0], timeout_ms=120_000)
        _same(first, want)
        P.server_stop()
        stopped = True
        th.join(timeout=60)
        assert not th.is_alive()
        # the blocked wait ends with the ticket's final status: its result, or
        # STATE for a query the stop left unadmitted
        assert "result" in outcome or outcome.get("code") == _lib.ERR_STATE, outcome
        if "result" in outcome:
            _same(outcome["result"], want)
        # every other ticket still reports its own final status after the stop
        done = failed = 0
        for t in tickets[1:-1]:
            try:
                _same(P.server_wait(t, timeout_ms=0), want)
                done += 1
            except _lib.SvmError as e:
                assert e.code == _lib.ERR_STATE, e
                failed += 1
        assert done + failed == len(tickets) - 2
        with pytest.raises(_lib.SvmError) as e:
            P.server_wait(tickets[1], timeout_ms=0)  # already claimed
        assert e.value.code in (_lib.ERR_NOT_FOUND, _lib.ERR_STATE)
    finally:
        if not stopped:
            P.server_stop()


def test_select_counts_pending_until_read(engine):
    """A dimension select whose result the caller does not take returns with its row
    count still being computed on the device; looking the table up (its row count, its
    columns, a view, the fact pipeline that joins it) waits for the count, and dropping
    the table right away waits for its kernels before its columns are freed."""
    wl = ssb.four_way(seed=41, fact_rows=2_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000,
                      prefix="pend")
    ssb.load_device(wl)
    tabs = host_tables(wl)
    want, dims = oracle_query(wl, tabs)
    # dropped while pending, several times over: each drop waits for its launch
    for _ in range(3):
        for pt, src, refs in wl.dims:
            plan = pt.plan(src, refs)
            P.run_plan(plan, want_result=False)
            P.table_drop(plan["id"])
    # each select's table, looked up while its count may still be pending
    for pt, src, refs in wl.dims:
        plan = pt.plan(src, refs)
        P.run_plan(plan, want_result=False)
        w = dims[plan["id"]]
        rows, _ = P.table_info(plan["id"])
        assert rows == w.rows, (plan["id"], rows, w.rows)
        assert_same_columns([P.table_column(plan["id"], c) for c in range(len(w.columns))], w.columns)
    # a view of a freshly selected table, then the fact pipeline over pending objects
    pt, src, refs = wl.dims[1]
    plan = pt.plan(src, refs)
    P.run_plan(plan, want_result=False)
    P.table_view("pend_view", plan["id"], 0, 1)
    assert P.table_info("pend_view")[0] == 1
    P.table_drop("pend_view")
    ssb.run_dims(wl)
    _same(ssb.run_fact(wl), want)
    for t in wl.tables:
        P.table_drop(t.id)
