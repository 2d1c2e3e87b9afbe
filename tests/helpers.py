"""Shared helpers: host copies of workload tables and oracle runs of their plans."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
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
