"""The host engine and hybrid CPU/GPU execution on the GPU box, against the oracle.

* PipelineInfo.isGPU = false (service/service.cu:51-52 -> PipelineTest::Run(...,
  isGPU), test/test.hpp:122-146): the dimension pipelines and the fact pipeline run
  their per-row transform on host cores, from code generated out of the same
  lowered plan as the kernels; dimension objects and results equal the oracle's.
* Hybrid execution (ResMgr::enable_hybrid_exec_ with CPU_ratio_ 3,
  core/pipeline.cu:172-180): bulk i of the fact pipeline runs on the host when
  i % 10 < 3 and on the GPU otherwise, concurrently; the host's partial scalar sums,
  group tables (private per task, or one shared table) and wide-group entries are
  merged with the device's.
"""
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query
from .test_gpu_wide_groups import _regroup

pytestmark = pytest.mark.gpu


def _same(got, want):
    assert got.kind == want.kind
    assert got.rows == want.rows, (got.rows, want.rows)
    assert_same_columns(got.columns, want.columns)


def _run_host(wl):
    """Every pipeline of the workload with isGPU = false; returns (dims, result)."""
    dims = {}
    for pt, src, refs in wl.dims:
        dims[pt.pipe_info["id"]] = pt.Run(src, refs, isGPU=False)
    fq, src, refs = wl.fact
    return dims, fq.Run(src, refs, isGPU=False)


@pytest.mark.parametrize("name", ["single_join", "three_way", "four_way"])
def test_is_gpu_false_runs_on_the_host_engine(engine, name):
    wl = {
        "single_join": lambda: ssb.single_join(seed=81, fact_rows=4_000_001, prefix="hs"),
        "three_way": lambda: ssb.three_way(seed=82, fact_rows=4_000_003, n_part=100_000, n_supp=10_000, prefix="ht"),
        "four_way": lambda: ssb.four_way(seed=83, fact_rows=4_000_005, n_cust=300_000, n_supp=20_000,
                                         n_part=140_000, prefix="hf"),
    }[name]()
    tabs = host_tables(wl)
    want, want_dims = oracle_query(wl, tabs)
    ssb.load_device(wl)
    dims, got = _run_host(wl)
    for pid, d in dims.items():
        print(f"[hybrid] {name}: host-engine dimension {pid}: {d.rows} rows", flush=True)
        _same(d, want_dims[pid])
    print(f"[hybrid] {name}: host-engine result {got.rows} rows", flush=True)
    _same(got, want)
    # the same plans on the GPU answer alike
    _same(ssb.run_query(wl), want)
    for t in wl.tables:
        P.table_drop(t.id)


@pytest.fixture(scope="module")
def hybrid_four_way(engine):
    # 11 bulks of 3M rows: bulks 0, 1, 2 and 10 on the host, 3..9 on the GPU
    wl = ssb.four_way(seed=85, fact_rows=33_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000, prefix="hy")
    tabs = host_tables(wl, threads=16)
    _, dims = oracle_query(wl, tabs, threads=16)
    ssb.load_device(wl)
    ssb.run_dims(wl)
    P.set_hybrid(True, 3, 16)
    yield wl, tabs, dims
    P.set_hybrid(False)


@pytest.mark.parametrize("case", ["standard", "shared_table", "wide", "wide_host_only"])
def test_hybrid_split_matches_oracle(hybrid_four_way, case):
    wl, tabs, dims = hybrid_four_way
    fq, src, refs = wl.fact
    if case == "standard":   # (d_year, c_city, s_city): per-task group tables
        pt = fq
    elif case == "shared_table":  # (c_city, suppkey): 5M groups, one shared host table
        pt, src, refs = _regroup(wl, "hy_shared", "$(2.1), $(0.2)")
    else:  # (custkey, suppkey, d_year): packed keys, host entries join the device's
        pt, src, refs = _regroup(wl, "hy_wide", "$(0.1), $(0.2), $(1.1)")
    want = O.run_pipeline(pt.plan(src, refs), tabs[src], [dims[r].table for r in refs], threads=16)
    if case == "wide_host_only":
        got = pt.Run(src, refs, isGPU=False)
    else:
        got = pt.Run(src, refs)
    print(f"[hybrid] {case}: {got.rows} groups", flush=True)
    assert got.rows > 0
    _same(got, want)


def test_hybrid_scalar(engine):
    wl = ssb.single_join(seed=86, fact_rows=33_000_001, prefix="hsc")
    tabs = host_tables(wl, threads=16)
    want, _ = oracle_query(wl, tabs, threads=16)
    ssb.load_device(wl)
    P.set_hybrid(True, 3, 16)
    try:
        _same(ssb.run_query(wl), want)
    finally:
        P.set_hybrid(False)
    for t in wl.tables:
        P.table_drop(t.id)
