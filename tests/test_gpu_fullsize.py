"""GPU parity at the workloads' full sizes, bit-exact against the CPU oracle.

  single join : date 2,556 rows, fact 6,000,000 rows, scalar SUM(extendedprice*discount)
  three-way   : part 800,000, supplier 20,000, date 2,556, fact 60,000,000 rows,
                group (d_year, p_brand), SUM(revenue)
  four-way    : customer 3,000,000, supplier 200,000, part 1,400,000, date 2,556,
                fact 600,000,000 rows (Zipf 1.1 foreign keys), group
                (d_year, c_city, s_city), SUM(revenue - supplycost)

  batched     : 64 template queries over slices of 1M-60M rows of the four-way fact table
  serving     : the 90/10 light/heavy mix with Poisson arrivals at 500 queries/s, heavy
                queries over all 600M rows, continuous batching and the plain path

The host copy of every table is generated with the same seeded generator as the
HBM copy (byte-identical), the oracle runs the plans on it, and every output
byte is compared: the dimension objects, then the group keys and int64 sums in
key order (or the scalar row).  The four-way tables need about 15 GB of host
memory and 15 GB of HBM.
"""
import time

import pytest

from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query, oracle_slice_queries

pytestmark = pytest.mark.gpu


def _full(engine, wl, threads=16):
    t0 = time.perf_counter()
    print(f"[fullsize] {wl.name} fact_rows={wl.fact_rows} constants={wl.constants}", flush=True)
    tabs = host_tables(wl, threads=threads)
    print(f"[fullsize] host tables {time.perf_counter() - t0:.1f} s", flush=True)
    want, want_dims = oracle_query(wl, tabs, threads=threads)
    print(f"[fullsize] oracle {time.perf_counter() - t0:.1f} s: {want.rows} rows", flush=True)
    del tabs
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
    print(f"[fullsize] {wl.name}: {got.rows} rows bit-exact ({time.perf_counter() - t0:.1f} s)", flush=True)
    return got


@pytest.mark.parametrize("seed", [1, 2])
def test_single_join_6m(engine, seed):
    got = _full(engine, ssb.single_join(seed=seed, prefix=f"fs1_{seed}"))
    assert got.rows == 1 and got.columns[0][0] > 0


def test_three_way_60m(engine):
    got = _full(engine, ssb.three_way(seed=1, prefix="fs2"))
    # 7 years x 40 brands of the chosen category
    assert got.rows == 280


@pytest.fixture(scope="module")
def four_way_600m(engine):
    """The four-way tables at full size in HBM and on the host (same bytes), and the
    oracle's dimension objects; dropped after the module (14.4 GB of fact rows)."""
    wl = ssb.four_way(seed=1, prefix="fs3")
    t0 = time.perf_counter()
    tabs = host_tables(wl, threads=16)
    print(f"[fullsize] four-way host tables {time.perf_counter() - t0:.1f} s", flush=True)
    ssb.load_device(wl)
    yield wl, tabs
    from shadowvm_amd import pipeline as P
    for t in wl.tables:
        if P.table_exists(t.id):
            P.table_drop(t.id)


def test_four_way_600m(four_way_600m):
    wl, tabs = four_way_600m
    t0 = time.perf_counter()
    want, want_dims = oracle_query(wl, tabs, threads=16)
    print(f"[fullsize] oracle {time.perf_counter() - t0:.1f} s: {want.rows} rows", flush=True)
    for pt, src, refs in wl.dims:
        got = pt.Run(src, refs)
        w = want_dims[pt.pipe_info["id"]]
        assert got.rows == w.rows
        assert_same_columns(got.columns, w.columns)
    got = ssb.run_fact(wl)
    assert got.rows == want.rows and 1000 < got.rows < 100_000
    assert_same_columns(got.columns, want.columns)


def test_batch_of_64_at_size(four_way_600m):
    """The batched form at its size: 64 template queries (1-4 joins, fact slices of
    1M-60M rows), fused into shared scans, every result equal to the oracle's."""
    from shadowvm_amd import pipeline as P
    wl, tabs = four_way_600m
    qs = ssb.batch_queries(wl, n=64, seed=7, window=60_000_000, min_rows=1_000_000, prefix="fsb")
    ssb.make_views(qs)
    plans, pos = [], []
    for q in qs:
        plans += q.plans()
        pos.append(len(plans) - 1)
    res, fused = P.batch_run(plans)
    print(f"[fullsize] batch: {fused} of {len(qs)} fused", flush=True)
    assert fused >= 48
    t0 = time.perf_counter()
    want = oracle_slice_queries(qs, tabs, tabs[wl.fact[1]], workers=8, threads=2)
    print(f"[fullsize] batch oracle {time.perf_counter() - t0:.1f} s; groups {[w.rows for w in want]}", flush=True)
    for q, w, k in zip(qs, want, pos):
        got = res[k]
        assert got is not None and got.rows == w.rows, (q.qid, got and got.rows, w.rows)
        assert_same_columns(got.columns, w.columns)
    for q in qs:
        P.table_drop(q.view[0])


@pytest.mark.parametrize("mode", [0, 1])
def test_serving_mix_at_size(four_way_600m, mode):
    """The serving form at its size: the 90 % light / 10 % heavy mix (heavy = the
    four-way join over all 600M rows) submitted with open-loop Poisson arrivals at
    500 queries/s to continuous batching (mode 0) and to the plain path (mode 1);
    every answer equals the oracle's."""
    import random

    from shadowvm_amd import pipeline as P
    wl, tabs = four_way_600m
    qs = ssb.serve_queries(wl, 40, seed=11 + mode, prefix=f"fss{mode}_")
    ssb.make_views(qs)
    rng = random.Random(5 + mode)
    P.server_start(workers=2, budget=0, max_batch=64, mode=mode)
    try:
        tickets = []
        t_next = time.perf_counter()
        for q in qs:
            t_next += rng.expovariate(500.0)
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(P.server_submit(q.plans()))
        got = [P.server_wait(t, timeout_ms=300_000) for t in tickets]
    finally:
        P.server_stop()
    heavy = sum(q.shape["kind"] == "heavy" for q in qs)
    print(f"[fullsize] serving mode {mode}: {len(qs)} queries ({heavy} heavy) answered", flush=True)
    t0 = time.perf_counter()
    want = oracle_slice_queries(qs, tabs, tabs[wl.fact[1]], workers=8, threads=2)
    print(f"[fullsize] serving oracle {time.perf_counter() - t0:.1f} s", flush=True)
    for q, g, w in zip(qs, got, want):
        assert g.rows == w.rows, (q.qid, q.shape, g.rows, w.rows)
        assert_same_columns(g.columns, w.columns)
    for q in qs:
        P.table_drop(q.view[0])
