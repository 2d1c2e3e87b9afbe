"""Pins the oracle's query semantics against textbook relational brute force
(numpy / dict joins, no hashing) on small seeded tables, and its bulk structure
against the merge invariant: results do not depend on tuple_per_bulk."""
import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import ssb

from .helpers import host_tables, oracle_query


def _hash31(b):
    h = 0
    for x in bytes(b):
        h = (h * 31 + (x - 256 if x > 127 else x)) & 0xFFFFFFFF
    return h & 0x7FFFFFFF


def _group(keys, vals):
    out = {}
    for k, v in zip(zip(*keys), vals):
        out[k] = out.get(k, 0) + int(v)
    ks = sorted(out)
    return [np.array([k[q] for k in ks], np.int32) for q in range(len(keys))] + \
        [np.array([out[k] for k in ks], np.int64)]


def brute_single_join(wl, tabs):
    c = wl.constants
    date = tabs[wl.tables[0].id]
    od, qty, disc, price = tabs[wl.tables[1].id].columns
    year_of = dict(zip(date.columns[0].tolist(), date.columns[1].tolist()))
    ok = np.array([year_of.get(k) == c["year"] for k in od.tolist()], bool)
    m = ok & (disc >= c["discount"]) & (disc <= c["discount"] + 2) & (qty < c["quantity"])
    return [np.array([int((price[m].astype(np.int64) * disc[m]).sum())], np.int64)]


def brute_three_way(wl, tabs):
    c = wl.constants
    date, part, supp, fact = (tabs[t.id] for t in wl.tables)
    year_of = dict(zip(date.columns[0].tolist(), date.columns[1].tolist()))
    brand_of = {k: b for k, cat, b in zip(*[x.tolist() for x in part.columns]) if cat == c["category"]}
    supp_ok = {k for k, r in zip(supp.columns[0].tolist(), supp.columns[1].tolist()) if r == c["region"]}
    od, pk, sk, rev = (x.tolist() for x in fact.columns)
    ky, kb, v = [], [], []
    for a, b, s, r in zip(od, pk, sk, rev):
        if a in year_of and b in brand_of and s in supp_ok:
            ky.append(year_of[a]); kb.append(brand_of[b]); v.append(r)
    return _group([ky, kb], v)


def brute_four_way(wl, tabs):
    c = wl.constants
    date, cust, supp, part, fact = (tabs[t.id] for t in wl.tables)
    year_of = {k: y for k, y in zip(date.columns[0].tolist(), date.columns[1].tolist())
               if c["year"] <= y <= c["year"] + 1}

    def city_map(t, nation):
        want = f"NATION_{nation:02d}".encode()
        return {int(k): _hash31(city) for k, city, nat in zip(t.columns[0], t.columns[1], t.columns[2])
                if bytes(nat).rstrip(b"\0") == want}
    ccity, scity = city_map(cust, c["c_nation"]), city_map(supp, c["s_nation"])
    parts = set(part.columns[0].tolist())
    od, ck, sk, pk, rev, cost = (x.tolist() for x in fact.columns)
    k1, k2, k3, v = [], [], [], []
    for a, b, s, p, r, q in zip(od, ck, sk, pk, rev, cost):
        if a in year_of and b in ccity and s in scity and p in parts:
            k1.append(year_of[a]); k2.append(ccity[b]); k3.append(scity[s]); v.append(r - q)
    return _group([k1, k2, k3], v)


CASES = [
    (lambda: ssb.single_join(seed=4, fact_rows=120_000, prefix="b1"), brute_single_join),
    (lambda: ssb.three_way(seed=6, fact_rows=150_000, n_part=20_000, n_supp=1_000, prefix="b2"),
     brute_three_way),
    (lambda: ssb.four_way(seed=8, fact_rows=200_000, n_cust=20_000, n_supp=2_000, n_part=9_000,
                          prefix="b3"), brute_four_way),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_equals_brute_force(case):
    make, brute = CASES[case]
    wl = make()
    tabs = host_tables(wl)
    got, dims = oracle_query(wl, tabs, tuple_per_bulk=3_000_000)
    want = brute(wl, tabs)
    assert all(v == 0 for v in got.dropped.values())
    assert len(got.columns) == len(want)
    for g, w in zip(got.columns, want):
        assert np.array_equal(g, w)
    if case > 0:
        assert got.rows > 5


@pytest.mark.parametrize("tpb", [7_777, 50_000])
def test_bulk_merge_invariant(tpb):
    """Per-bulk sort + reduce then the cross-bulk merge equals one bulk."""
    wl = ssb.four_way(seed=9, fact_rows=200_000, n_cust=20_000, n_supp=2_000, n_part=9_000, prefix="m3")
    tabs = host_tables(wl)
    one, _ = oracle_query(wl, tabs, tuple_per_bulk=3_000_000)
    many, _ = oracle_query(wl, tabs, tuple_per_bulk=tpb, dim_tuple_per_bulk=3_000_000)
    assert many.rows == one.rows > 5
    for a, b in zip(one.columns, many.columns):
        assert np.array_equal(a, b)


def test_dimension_select_keeps_row_order():
    wl = ssb.four_way(seed=10, fact_rows=1000, n_cust=5_000, n_supp=500, n_part=100, prefix="o3")
    tabs = host_tables(wl)
    _, dims = oracle_query(wl, tabs)
    cust = tabs[wl.tables[1].id]
    want = f"NATION_{wl.constants['c_nation']:02d}".encode()
    keep = [r for r in range(cust.rows) if bytes(cust.columns[2][r]).rstrip(b"\0") == want]
    d = dims[wl.dims[1][0].pipe_info["id"]]
    assert d.rows == len(keep)
    assert np.array_equal(d.columns[0], cust.columns[0][keep])
    assert np.array_equal(d.columns[1], cust.columns[1][keep])


def test_scalar_closed_form_all_rows_pass():
    """With predicates that every row satisfies, SUM(price*discount) is the plain
    dot product over the whole table."""
    wl = ssb.single_join(seed=2, fact_rows=50_000, prefix="cf")
    tabs = host_tables(wl)
    fq, src, refs = wl.fact
    plan = fq.plan(src, refs)
    plan["workflow"][1]["args"]["predication"] = "$(0.1)>=1 && $(0.2)>=0"
    date = tabs[wl.tables[0].id]
    dq = O.run_pipeline({"id": "all_dates", "workflow": [
        {"type": 0, "args": {"type": "object"}, "output": None},
        {"type": 3, "args": {"select": "$(0.0), $(0.1)"}, "output": None},
        {"type": 7, "args": {}, "output": {"column_list": [{"type": 2, "val_size": 4}] * 2 +
                                           [{"type": 7, "val_size": 40}]}}]}, date)
    res = O.run_pipeline(plan, tabs[src], [dq.table])
    _, _, disc, price = tabs[src].columns
    assert res.columns[0][0] == int((price.astype(np.int64) * disc).sum())


def test_group_exists_even_when_sum_is_zero():
    """A group exists iff a row passes, even when its sum is 0 (discount 0 rows)."""
    date = O.OTable([2, 2, 2], [4, 4, 4], [np.array([1, 2], np.int32), np.array([10, 20], np.int32),
                                            np.array([1, 1], np.int32)])
    fact = O.OTable([2, 2], [4, 4], [np.array([1, 2, 2, 3], np.int32), np.array([0, 5, -5, 9], np.int32)])
    dq = O.run_pipeline({"id": "d", "workflow": [
        {"type": 0, "args": {"type": "object"}},
        {"type": 3, "args": {"select": "$(0.0), $(0.1)"}},
        {"type": 7, "args": {}, "output": {"column_list": [{"type": 2, "val_size": 4}] * 2 +
                                           [{"type": 7, "val_size": 40}]}}]}, date)
    fq = {"id": "f", "workflow": [
        {"type": 0, "args": {"type": "object"}},
        {"type": 6, "args": {"lkey": "$(0.0)", "rkey": "$(1.0)"}},
        {"type": 4, "args": {"dim": "$(1.1)", "metric": "$(0.1)"}},
        {"type": 7, "args": {}, "output": {"column_list": [{"type": 2, "val_size": 4}] * 3}}]}
    r = O.run_pipeline(fq, fact, [dq.table])
    assert r.columns[0].tolist() == [10, 20]
    assert r.columns[1].tolist() == [0, 0]


def wrapping_plan():
    """Group by a source column; the metric's int32 product overflows on most rows."""
    from shadowvm_amd import pipeline as P
    I = P.INT32
    s = P.create_schema([I, I, I], [4, 4, 4])
    out = P.create_schema([I, I], [4, 4])
    pt = P.PipelineTest("wrap")
    pt.AddOperator(P.SOURCE, {"type": "object"}, s)
    pt.AddOperator(P.FILTER, {"predication": "$(0.1) != 7"}, out)
    pt.AddOperator(P.REDUCEBY, {"dim": "$(0.0)", "metric": "$(0.1)*$(0.2)+$(0.2)"}, out)
    pt.AddOperator(P.SINK, {}, out)
    return pt


def wrapping_table(n=50_000, seed=5):
    rng = np.random.default_rng(seed)
    k = rng.integers(-3, 4, n, dtype=np.int32)
    a = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    b = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    return k, a, b


def test_int32_metric_wraps():
    """Per-row int32 arithmetic wraps in two's complement (both sides compile with
    -fwrapv); only the accumulator is widened to int64."""
    k, a, b = wrapping_table()
    t = O.OTable([2, 2, 2], [4, 4, 4], [k, a, b], max_tuple_count=3_000_000)
    got = O.run_pipeline(wrapping_plan().plan("src", []), t, [])
    m = a != 7
    with np.errstate(over="ignore"):
        v = (a[m] * b[m] + b[m]).astype(np.int32)  # numpy int32 arithmetic wraps
    want = _group([k[m]], v.astype(np.int64))
    assert got.rows == len(want[0])
    for g, w in zip(got.columns, want):
        assert np.array_equal(g, w)


def test_zero_copy_view_equals_copied_slice():
    """OTable.view (a slice sharing its table's buffers) answers exactly like a copied
    slice, for a batch-template query over an odd row range."""
    wl = ssb.four_way(seed=3, fact_rows=200_003, n_cust=20_000, n_supp=2_000, n_part=14_000, prefix="ov")
    tabs = host_tables(wl, threads=4)
    q = ssb.batch_queries(wl, n=1, seed=2, window=200_003, min_rows=50_000, prefix="ovq")[0]
    fact = tabs[wl.fact[1]]
    _, _, row0, rows = q.view
    copied = O.OTable(fact.types, fact.sizes, [c[row0:row0 + rows] for c in fact.columns])
    dims = {pt.pipe_info["id"]: O.run_pipeline(pt.plan(src, refs), tabs[src], []) for pt, src, refs in q.dims}
    fq, src, refs = q.fact
    a = O.run_pipeline(fq.plan(src, refs), fact.view(row0, rows), [dims[r].table for r in refs])
    b = O.run_pipeline(fq.plan(src, refs), copied, [dims[r].table for r in refs])
    assert a.rows == b.rows
    for x, y in zip(a.columns, b.columns):
        assert np.array_equal(x, y)
