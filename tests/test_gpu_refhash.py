"""Joins on the GPU reproduce the reference's hash tables exactly: inserts dropped by
the 10-probe table (com/cpp/cpu-runtime.hpp:80-91) never join, and only bulk 0 of a
dimension is read (core/pipeline.cu:91,161-162,294-296), on the plain path, in a
batch (queries whose dimension objects drop go to the plain path) and in the server.
The engine is restarted with other bulk capacities (tuple_per_bulk) for some cases."""
import contextlib

import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns, edge_query, host_tables, oracle_query
from .test_oracle_bulks import I, bulk0_case, bulk0_expected, colliding_case, dim_plan, fact_plan

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def engine_with(engine, tpb):
    print(f"[refhash] engine tuple_per_bulk={tpb}", flush=True)
    engine.shutdown()
    engine.init(device=0, tuple_per_bulk=tpb)
    try:
        yield engine
    finally:
        engine.shutdown()
        engine.init(device=0, tuple_per_bulk=3_000_000)


# Dropped inserts of primary keys 1..N inserted in order: the counts an independent
# simulation of the reference's insertion found (also the oracle's known answers).
@pytest.mark.parametrize("n,slots,want", [(1_400_000, 10_000_000, 6910), (800_000, 10_000_000, 630),
                                          (1_400_000, 15_000_000, 0), (3_000_000, 15_000_000, 746_032),
                                          (2556, 10_000_000, 0)])
def test_published_drop_counts(engine, n, slots, want):
    P.table_upload("rh_keys", [I], [4], [np.arange(1, n + 1, dtype=np.int32)])
    try:
        assert P.hash_drops("rh_keys", 0, slots) == want
    finally:
        P.table_drop("rh_keys")


def test_golden_drop_counts(engine):
    """Dropped inserts computed on the GPU equal those of the reference's own runtime
    (golden vectors from oracle/_ref): sparse and crowded tables, duplicate keys
    (every row takes a slot) and negative keys."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "runtime_vectors.json")) as f:
        tables = json.load(f)["int_tables"]
    for q, t in enumerate(tables):
        print(f"[refhash] golden table {q}: {len(t['keys'])} keys, {t['slots']} slots", flush=True)
        P.table_upload("rh_golden", [I], [4], [np.array(t["keys"], np.int32)])
        try:
            assert P.hash_drops("rh_golden", 0, t["slots"]) == t["dropped"], q
        finally:
            P.table_drop("rh_golden")


def test_tiny_bulk0_and_drops(engine):
    """The oracle's brute-force cases through the product, at a bulk capacity of 16."""
    with engine_with(engine, 16):
        base, fact = bulk0_case()
        P.table_upload("tb_base", [I, I], [4, 4], base)
        P.table_upload("tb_fact", [I, I], [4, 4], fact)
        dim_plan("tb_dim", "$(0.1)==1").Run("tb_base", want_result=False)
        got = fact_plan("tb_q", "", "$(0.1)").Run("tb_fact", ["tb_dim"])
        assert int(got.columns[0][0]) == bulk0_expected(base, fact, 16)[1]

        base, fact, kept = colliding_case()
        P.table_upload("tc_base", [I, I], [4, 4], base)
        P.table_upload("tc_fact", [I, I], [4, 4], fact)
        dim_plan("tc_dim", "").Run("tc_base", want_result=False)
        assert P.hash_drops("tc_dim", 0) == len(base[0]) - 10
        want_sum = int(sum(int(v) for k, v in zip(*fact) if int(k) in kept))
        got = fact_plan("tc_q", "", "$(0.1)").Run("tc_fact", ["tc_dim"])
        assert int(got.columns[0][0]) == want_sum
        # the uploaded table itself as the dimension (its own kept rows, not a subset's)
        got = fact_plan("tc_q2", "", "$(0.1)").Run("tc_fact", ["tc_base"])
        assert int(got.columns[0][0]) == want_sum
        g = fact_plan("tc_g", "$(1.1)", "$(0.1)+$(1.0)").Run("tc_fact", ["tc_dim"])
        want = O.run_pipeline(fact_plan("tc_g", "$(1.1)", "$(0.1)+$(1.0)").plan("f", ["d"]),
                              O.OTable([I, I], [4, 4], fact, 16),
                              [O.run_pipeline(dim_plan("tc_dim", "").plan("b"), O.OTable([I, I], [4, 4], base, 16),
                                              tuple_per_bulk=16).table], tuple_per_bulk=16)
        assert g.rows == want.rows == 1
        assert_same_columns(g.columns, want.columns)


def test_four_way_at_the_reference_bulk_size(engine):
    """tuple_per_bulk = 2,000,000 (the reference's default, com/cpp/com.cu:13): the
    unfiltered 1.4M-row part dimension drops 6,910 inserts in its 10M-slot table and
    the 3M-row customer table spans two bulks, so its dimension joins only the rows
    of the first.  Both shape the answer; the GPU matches the oracle bit for bit."""
    tpb = 2_000_000
    with engine_with(engine, tpb):
        wl = ssb.four_way(seed=41, fact_rows=3_000_000, prefix="r2")
        tabs = host_tables(wl, tuple_per_bulk=tpb)
        want, dims = oracle_query(wl, tabs, tuple_per_bulk=tpb)
        assert want.dropped[4] == 6910
        cust = dims[wl.dims[1][0].pipe_info["id"]].table
        assert cust.bulk0_rows < cust.rows
        ssb.load_device(wl)
        got = ssb.run_query(wl)
        print(f"[refhash] four-way at 2M bulks: {got.rows} groups", flush=True)
        assert got.rows == want.rows > 10
        assert_same_columns(got.columns, want.columns)
        assert P.hash_drops(wl.dims[3][0].pipe_info["id"], 0) == 6910


def _oracle_of(q, tabs, fact):
    vid, base, row0, rows = q.view
    view = O.OTable(fact.types, fact.sizes, [c[row0:row0 + rows] for c in fact.columns], max_tuple_count=3_000_000)
    dims = {}
    for pt, src, refs in q.dims:
        dims[pt.pipe_info["id"]] = O.run_pipeline(pt.plan(src, refs), tabs[src], [])
    fq, src, refs = q.fact
    return O.run_pipeline(fq.plan(src, refs), view, [dims[r].table for r in refs])


def test_batch_and_server_leave_dropping_queries_to_the_plain_path(engine):
    """With the 3M-row customer table the 15M-slot table drops 746,032 of its inserts:
    a batched query whose customer dimension keeps 96% of the rows drops too and is
    answered on the plain path, one keeping 4% drops nothing and stays fused."""
    wl = ssb.four_way(seed=43, fact_rows=2_000_000, prefix="rr")
    ssb.load_device(wl)
    n = wl.fact_rows
    qs = [
        edge_query(wl, "rr0", 0, n, [("customer", 'StrCmp($(0.2),"NATION_03",10)!=0', "$(0.0), $(0.3)", 1)],
                   "", "$(1.1)", "$(0.4)"),
        edge_query(wl, "rr1", 0, n, [("customer", 'StrCmp($(0.2),"NATION_03",10)==0', "$(0.0), $(0.3)", 1)],
                   "", "$(1.1)", "$(0.4)"),
        edge_query(wl, "rr2", 5, n - 9,
                   [("date", "$(0.1)==1994", "$(0.0), $(0.1)", 0),
                    ("customer", 'StrCmp($(0.3),"REGION_1",8)==0', "$(0.0), $(0.1)", 1)],
                   "", "$(1.1), $(2.1)", "$(0.4)-$(0.5)"),
        edge_query(wl, "rr3", 0, n, [("part", "", "$(0.0), $(0.2)", 3)], "", "$(1.1)", "$(0.5)"),
    ]
    ssb.make_views(qs)
    tabs = host_tables(wl)
    fact = tabs[wl.fact[1]]
    want = [_oracle_of(q, tabs, fact) for q in qs]
    drops = [sum(w.dropped.values()) for w in want]
    print(f"[refhash] oracle drops per query: {drops}", flush=True)
    assert drops[0] > 0 and drops[1] == 0 and drops[3] == 0
    plans, pos = [], []
    for q in qs:
        plans += q.plans()
        pos.append(len(plans) - 1)
    res, fused = P.batch_run(plans)
    print(f"[refhash] batch fused {fused} of {len(qs)}", flush=True)
    assert fused == sum(1 for d in drops if d == 0)
    for q, k, w in zip(qs, pos, want):
        assert res[k].rows == w.rows, q.qid
        assert_same_columns(res[k].columns, w.columns)
    P.server_start(workers=2, max_batch=8, mode=0)
    try:
        tickets = [P.server_submit(q.plans()) for q in qs]
        for q, t, w in zip(qs, tickets, want):
            r = P.server_wait(t, 60_000)
            assert r.rows == w.rows, q.qid
            assert_same_columns(r.columns, w.columns)
    finally:
        P.server_stop()


@pytest.mark.parametrize("dups", [0, 1, 3000])
def test_dimension_with_repeated_keys(engine, dups):
    """A dimension root whose key column repeats keys (`dups` extra rows drawn from its
    keys, shuffled in): its filtered object is not built as a one-row-per-key map, and
    every probe finds the first row of its key, as in the oracle.  With dups == 0 the
    keys are unique and the object takes the direct build."""
    rng = np.random.default_rng(40 + dups)
    keys = np.arange(1, 20_001, dtype=np.int32)
    keys = rng.permutation(np.concatenate([keys, rng.choice(keys, dups)])).astype(np.int32)
    attr = (np.arange(len(keys)) % 97).astype(np.int32)  # differs between a key's rows
    base = [keys, attr]
    fk = rng.integers(1, 20_100, 1_000_003).astype(np.int32)
    val = rng.integers(-1000, 1000, len(fk)).astype(np.int32)
    fact = [fk, val]
    tag = f"rk{dups}"
    print(f"[refhash] repeated keys: {dups} extra rows", flush=True)
    P.table_upload(f"{tag}_base", [I, I], [4, 4], base)
    P.table_upload(f"{tag}_fact", [I, I], [4, 4], fact)
    try:
        dim_plan(f"{tag}_dim", "$(0.1)!=5").Run(f"{tag}_base", want_result=False)
        q = fact_plan(f"{tag}_q", "$(1.1)", "$(0.1)")
        got = q.Run(f"{tag}_fact", [f"{tag}_dim"])
        d = O.run_pipeline(dim_plan(f"{tag}_dim", "$(0.1)!=5").plan("b"), O.OTable([I, I], [4, 4], base, 3_000_000),
                           tuple_per_bulk=3_000_000)
        want = O.run_pipeline(q.plan("f", ["d"]), O.OTable([I, I], [4, 4], fact, 3_000_000), [d.table],
                              tuple_per_bulk=3_000_000)
        assert got.rows == want.rows == 96
        assert_same_columns(got.columns, want.columns)
    finally:
        for t in (f"{tag}_base", f"{tag}_fact", f"{tag}_dim"):
            P.table_drop(t)
