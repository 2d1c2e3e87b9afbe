"""Group-by over key spaces too wide for a group table, on the GPU against the oracle.

The reference sorts and reduces any int32 keys into any number of groups
(core/pipeline.cu:380-471).  Group spaces beyond a dense table (more than 2^27
code combinations, or a source-column key whose range exceeds 2^24) aggregate
through packed 64-bit keys, LDS pre-aggregation, HBM append buffers and a sort +
reduce by key.  Every output byte is compared with the oracle: group keys in
lexicographic signed order and int64 sums.
"""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query

pytestmark = pytest.mark.gpu


def _regroup(wl, qid, dim, metric="$(0.4)-$(0.5)", joins=None):
    """The workload's fact pipeline under a new id with another ReduceBy key list
    (and optionally only some of its joins, given as ref positions 1..4)."""
    fq, src, refs = wl.fact
    pt = P.PipelineTest(qid)
    pt.pipe_info = copy.deepcopy(fq.pipe_info)
    pt.pipe_info["id"] = qid
    wf = []
    keep = [r for k, r in enumerate(refs) if joins is None or k + 1 in joins]
    for op in pt.pipe_info["workflow"]:
        op["id"] = qid
        if op["type"] == P.JOINWITH and joins is not None:
            k = int(op["args"]["rkey"][2])
            if k not in joins:
                continue
            op["args"]["rkey"] = f"$({joins.index(k) + 1}.0)"
        if op["type"] == P.REDUCEBY:
            op["args"] = {"dim": dim, "metric": metric}
        wf.append(op)
    pt.pipe_info["workflow"] = wf
    return pt, src, keep


@pytest.fixture(scope="module")
def four_way_wide(engine):
    wl = ssb.four_way(seed=71, fact_rows=4_000_000, n_cust=3_000_000, n_supp=200_000, n_part=140_000,
                      prefix="wg")
    tabs = host_tables(wl)
    _, dims = oracle_query(wl, tabs)
    ssb.load_device(wl)
    ssb.run_dims(wl)
    return wl, tabs, dims


def _check(pt, src, refs, tabs, dims, src_table=None):
    want = O.run_pipeline(pt.plan(src, refs), src_table or tabs[src], [dims[r].table for r in refs])
    got = pt.Run(src, refs)
    print(f"[wide] {pt.pipe_info['id']}: {got.rows} groups", flush=True)
    assert got.rows == want.rows, (got.rows, want.rows)
    assert_same_columns(got.columns, want.columns)
    return got


def test_customer_supplier_year(four_way_wide):
    """(c_custkey, s_suppkey, d_year): 3M x 200K x 7 codes, far beyond a group table."""
    wl, tabs, dims = four_way_wide
    pt, src, refs = _regroup(wl, "wg_csy", "$(0.1), $(0.2), $(1.1)")
    got = _check(pt, src, refs, tabs, dims)
    assert got.rows > 100


def test_every_row_its_own_group(four_way_wide):
    """(custkey, suppkey, partkey) over all rows joined only with the unfiltered part
    object: nearly every row is a group of its own, so most rows miss the LDS table
    and go straight to the HBM entries."""
    wl, tabs, dims = four_way_wide
    pt, src, refs = _regroup(wl, "wg_all", "$(0.1), $(0.2), $(0.3)", metric="$(0.4)", joins=[4])
    got = _check(pt, src, refs, tabs, dims)
    assert got.rows > 1_000_000


def test_wide_source_key_and_overflowing_first_pass(engine):
    """A source-column key spanning the whole int32 range (32 bits of code), with a
    second key, over 24M rows that are nearly all distinct groups: more entries than
    the first pass's buffer holds, so the scan runs again with the bound."""
    from .test_oracle_brute import wrapping_table
    n = 24_000_000
    k, a, b = wrapping_table(n=n, seed=17)
    small = (np.arange(n, dtype=np.int64) % 5 - 2).astype(np.int32)
    P.table_upload("wg_src", [P.INT32] * 4, [4] * 4, [k, a, b, small])
    I = P.INT32
    s = P.create_schema([I] * 4, [4] * 4)
    out = P.create_schema([I, I, P.INT64, P.INT64], [4, 4, 8, 8])
    pt = P.PipelineTest("wg_wide")
    pt.AddOperator(P.SOURCE, {"type": "object"}, s)
    pt.AddOperator(P.FILTER, {"predication": "$(0.3) != 1"}, out)
    pt.AddOperator(P.REDUCEBY, {"dim": "$(0.1), $(0.3)", "metric": "$(0.0)*$(0.2), $(0.2)"}, out)
    pt.AddOperator(P.SINK, {}, out)
    src = O.OTable([2] * 4, [4] * 4, [k, a, b, small])
    want = O.run_pipeline(pt.plan("wg_src", []), src, [])
    got = pt.Run("wg_src", [])
    print(f"[wide] wg_wide: {got.rows} groups", flush=True)
    assert got.rows == want.rows and got.rows > 19_000_000
    assert_same_columns(got.columns, want.columns)
    P.table_drop("wg_src")
