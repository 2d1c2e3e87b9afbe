"""The reference's thirteen SSB query tests (test/test-SSB-Q11.cu ... Q43.cu) run
through libsvm on seeded SSB-shaped tables (1M lineorder rows, SF1 dimensions) and
compared with the oracle byte for byte: every dimension object, then the query's
group keys and int64 sums in key order (or its scalar)."""
import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P

from .helpers import assert_same_columns
from .ssb_queries import NAMES, gen_tables, queries

pytestmark = pytest.mark.gpu

IDS = {k: f"ssbq_{k}" for k in ("lineorder", "date", "customer", "supplier", "part")}


@pytest.fixture(scope="module")
def ssb_tables(engine):
    tabs = gen_tables(seed=7)
    for k, (types, sizes, cols) in tabs.items():
        if not P.table_exists(IDS[k]):
            P.table_upload(IDS[k], types, sizes, cols)
    host = {IDS[k]: O.OTable(types, sizes, cols, max_tuple_count=3_000_000) for k, (types, sizes, cols) in tabs.items()}
    return host


@pytest.mark.parametrize("name", NAMES)
def test_ssb_query_equals_oracle(ssb_tables, name):
    dims, fact = queries("ssbq_", IDS)[name]
    refs_host, ref_ids = [], []
    for pt, src in dims:
        want = O.run_pipeline(pt.plan(src), ssb_tables[src], [])
        got = pt.Run(src)
        assert got.rows == want.rows, (pt.pipe_info["id"], got.rows, want.rows)
        assert_same_columns(got.columns, want.columns)
        refs_host.append(want.table)
        ref_ids.append(pt.pipe_info["id"])
    want = O.run_pipeline(fact.plan(IDS["lineorder"], ref_ids), ssb_tables[IDS["lineorder"]], refs_host)
    got = fact.Run(IDS["lineorder"], ref_ids)
    print(f"[ssb] {name}: {got.rows} rows; dimension rows "
          f"{[r.rows for r in refs_host]}", flush=True)
    assert got.kind == want.kind
    assert got.rows == want.rows
    assert_same_columns(got.columns, want.columns)
    if got.kind == "group":
        assert got.rows > 0
