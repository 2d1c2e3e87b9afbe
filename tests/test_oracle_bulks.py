"""The oracle's bulk structure and hash-table drops against brute force on tiny
tables:

* a join reads only bulk 0 of its ref object (core/pipeline.cu:91,294-295), and a
  select's output bulk b holds what its source's bulk b produced
  (core/pipeline.cu:161-162,294-296), so a dimension built over a source of several
  bulks joins only the passing rows of the source's first bulk;
* the join's hash table has HashRate x bulk-capacity slots and drops an insert that
  finds no free slot in 10 probes (com/cpp/cpu-runtime.hpp:80-91); a dropped row
  never joins (com/cpp/cpu-runtime.hpp:125-138).

The expected answers are computed with sets and sums; the dropped rows follow from
the construction (every key shares one home slot), not from a second copy of the
insertion loop.
"""
import numpy as np

from oracle import oracle as O
from shadowvm_amd import pipeline as P

I, S = P.INT32, P.STRING


def hash31(v):
    h = 0
    for b in int(v & 0xFFFFFFFF).to_bytes(4, "little"):
        h = (h * 31 + (b - 256 if b > 127 else b)) & 0xFFFFFFFF
    return h & 0x7FFFFFFF


def dim_plan(pid, filt):
    pt = P.PipelineTest(pid)
    osch = P.create_schema([I, I, S], [4, 4, 40])
    pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I, I], [4, 4]))
    if filt:
        pt.AddOperator(P.FILTER, {"predication": filt}, osch)
    pt.AddOperator(P.MAP, {"select": "$(0.0), $(0.1)"}, osch)
    pt.AddOperator(P.SINK, {}, osch)
    return pt


def fact_plan(pid, dim, metric):
    """Scalar sum of `metric` over fact rows whose key $(0.0) joins the dimension
    (dim == ""), or grouped by the dimension's attribute $(1.1)."""
    pt = P.PipelineTest(pid)
    res = P.create_schema([I] * 4, [4] * 4)
    pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I, I], [4, 4]))
    pt.AddOperator(P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
    pt.AddOperator(P.REDUCEBY, {"dim": dim, "metric": metric}, res)
    pt.AddOperator(P.SINK, {}, res)
    return pt


def bulk0_case(seed=1, rows=40, fact_rows=300):
    """A (key, attr) base table with keys 1..rows and attr = key % 3, and a fact
    table of (key, value) rows; the dimension keeps attr == 1."""
    rng = np.random.default_rng(seed)
    keys = np.arange(1, rows + 1, dtype=np.int32)
    base = [keys, (keys % 3).astype(np.int32)]
    fk = rng.integers(1, rows + 6, fact_rows).astype(np.int32)
    val = rng.integers(-1000, 1000, fact_rows).astype(np.int32)
    return base, [fk, val]


def bulk0_expected(base, fact, tpb):
    keys, attr = base
    first = min(len(keys), tpb)
    joined = {int(k) for k, a in zip(keys[:first], attr[:first]) if a == 1}
    return joined, int(sum(int(v) for k, v in zip(*fact) if int(k) in joined))


def colliding_case(n=16, tpb=16, fact_rows=400, seed=2):
    """n keys that all hash to one home slot of the 5*tpb-slot table: the first 10
    inserted (in row order) take the home slot and the 9 after it, the rest drop."""
    slots = 5 * tpb
    home = hash31(1) % slots
    keys = [k for k in range(1, 200_000) if hash31(k) % slots == home][:n]
    assert len(keys) == n
    rng = np.random.default_rng(seed)
    keys = np.array(rng.permutation(keys), np.int32)
    base = [keys, np.ones(n, np.int32)]
    fk = rng.choice(np.concatenate([keys, np.array([7, 8], np.int32)]), fact_rows).astype(np.int32)
    val = rng.integers(0, 1000, fact_rows).astype(np.int32)
    kept = {int(k) for k in keys[:10]}
    return base, [fk, val], kept


def test_join_reads_bulk0_of_a_select_output():
    base, fact = bulk0_case()
    sums = {}
    for tpb in (16, 64):
        d = O.run_pipeline(dim_plan("ob_dim", "$(0.1)==1").plan("b"), O.OTable([I, I], [4, 4], base, tpb),
                           tuple_per_bulk=tpb)
        joined, want = bulk0_expected(base, fact, tpb)
        assert d.rows == sum(1 for a in base[1] if a == 1)
        assert d.table.bulk0_rows == len(joined)
        r = O.run_pipeline(fact_plan("ob_q", "", "$(0.1)").plan("f", ["ob_dim"]),
                           O.OTable([I, I], [4, 4], fact, tpb), [d.table], tuple_per_bulk=tpb)
        assert int(r.columns[0][0]) == want
        sums[tpb] = want
    assert sums[16] != sums[64]


def test_dropped_rows_never_join():
    base, fact, kept = colliding_case()
    tpb = 16
    d = O.run_pipeline(dim_plan("od_dim", "").plan("b"), O.OTable([I, I], [4, 4], base, tpb), tuple_per_bulk=tpb)
    r = O.run_pipeline(fact_plan("od_q", "", "$(0.1)").plan("f", ["od_dim"]),
                       O.OTable([I, I], [4, 4], fact, tpb), [d.table], tuple_per_bulk=tpb)
    assert r.dropped[1] == len(base[0]) - 10
    want = int(sum(int(v) for k, v in zip(*fact) if int(k) in kept))
    assert int(r.columns[0][0]) == want
    # grouped by the dimension's attribute through the same table
    g = O.run_pipeline(fact_plan("od_g", "$(1.1)", "$(0.1)+$(1.0)").plan("f", ["od_dim"]),
                       O.OTable([I, I], [4, 4], fact, tpb), [d.table], tuple_per_bulk=tpb)
    want_g = int(sum(int(v) + int(k) for k, v in zip(*fact) if int(k) in kept))
    assert g.rows == 1 and int(g.columns[0][0]) == 1 and int(g.columns[1][0]) == want_g
