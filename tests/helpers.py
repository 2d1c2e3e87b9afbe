"""Shared helpers: host copies of workload tables and oracle runs of their plans."""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np

from oracle import oracle as O
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb


def host_tables(wl, threads=8, tuple_per_bulk=3_000_000):
    tabs = {}
    for t in wl.tables:
        types, sizes, cols = ssb.host_table(wl, t, threads=threads)
        tabs[t.id] = O.OTable(types, sizes, cols, max_tuple_count=tuple_per_bulk)
    return tabs


def oracle_query(wl, tabs, tuple_per_bulk=3_000_000, threads=8, dim_tuple_per_bulk=None):
    """Run the workload's dimension pipelines then its fact pipeline on the oracle.
    Returns (fact result, {dim pipeline id: select result})."""
    dims = {}
    for pt, src, refs in wl.dims:
        r = O.run_pipeline(pt.plan(src, refs), tabs[src], [dims[x].table for x in refs],
                           tuple_per_bulk=dim_tuple_per_bulk or tuple_per_bulk, threads=threads)
        dims[pt.pipe_info["id"]] = r
    fq, src, refs = wl.fact
    res = O.run_pipeline(fq.plan(src, refs), tabs[src], [dims[x].table for x in refs],
                         tuple_per_bulk=tuple_per_bulk, threads=threads)
    return res, dims


def assert_same_columns(got, want):
    assert len(got) == len(want), (len(got), len(want))
    for q, (g, w) in enumerate(zip(got, want)):
        g, w = np.asarray(g), np.asarray(w)
        assert g.shape == w.shape, f"column {q}: shape {g.shape} != {w.shape}"
        if not np.array_equal(g, w):
            bad = np.nonzero(g != w)[0][:5] if g.ndim == 1 else np.nonzero((g != w).any(axis=1))[0][:5]
            raise AssertionError(f"column {q} differs at rows {bad.tolist()}: {g[bad]} vs {w[bad]}")


def edge_query(wl, qid, row0, rows, dims, fact_filter, keys, metric):
    """A fused-form query over the four-way tables: dims is a list of
    (dimension kind, filter, select, fact key column)."""
    I, S = P.INT32, P.STRING
    ids = {t.kind: t.id for t in wl.tables}
    schemas = {
        "date": (ids[ssb.GEN_DATE], [I, I, I], [4, 4, 4]),
        "customer": (ids[ssb.GEN_CUST3], [I, S, S, S], [4, 10, 10, 10]),
        "supplier": (ids[ssb.GEN_SUPP3], [I, S, S, S], [4, 10, 10, 10]),
        "part": (ids[ssb.GEN_PART3], [I, I, I], [4, 4, 4]),
    }
    view = (f"{qid}_slice", ids[ssb.GEN_FACT3], row0, rows)
    dpts, refs = [], []
    for k, (kind, filt, sel, _) in enumerate(dims):
        src, st, ss = schemas[kind]
        pt = P.PipelineTest(f"{qid}_d{k}")
        osch = P.create_schema([I, I, S], [4, 4, 40]) if kind in ("date", "part") else \
            P.create_schema([I, S, S], [4, 10, 40])
        pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema(st, ss))
        if filt:
            pt.AddOperator(P.FILTER, {"predication": filt}, osch)
        pt.AddOperator(P.MAP, {"select": sel}, osch)
        pt.AddOperator(P.SINK, {}, osch)
        dpts.append((pt, src, []))
        refs.append(pt.pipe_info["id"])
    res = P.create_schema([I] * 12, [4] * 12)
    fq = P.PipelineTest(qid)
    fq.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I] * 6, [4] * 6))
    if fact_filter:
        fq.AddOperator(P.FILTER, {"predication": fact_filter}, res)
    for k, (_, _, _, lcol) in enumerate(dims):
        fq.AddOperator(P.JOINWITH, {"lkey": f"$(0.{lcol})", "rkey": f"$({k + 1}.0)"}, res)
    fq.AddOperator(P.REDUCEBY, {"dim": keys, "metric": metric}, res)
    fq.AddOperator(P.SINK, {}, res)
    return ssb.Query(qid, view, dpts, (fq, view[0], refs), {})


def oracle_slice_query(q, tabs, fact, threads=4):
    """The oracle's answer to a batch-template query: its dimension pipelines over
    the dimension tables, then its fact pipeline over its slice of `fact`."""
    vid, base, row0, rows = q.view
    view = fact.view(row0, rows)
    dims = {}
    for pt, src, refs in q.dims:
        dims[pt.pipe_info["id"]] = O.run_pipeline(pt.plan(src, refs), tabs[src], [], threads=threads)
    fq, src, refs = q.fact
    return O.run_pipeline(fq.plan(src, refs), view, [dims[r].table for r in refs], threads=threads)


def oracle_slice_queries(qs, tabs, fact, workers=8, threads=4):
    """oracle_slice_query for every query, several at once (the plans compile in parallel)."""
    with cf.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda q: oracle_slice_query(q, tabs, fact, threads=threads), qs))
