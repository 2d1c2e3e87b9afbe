"""svm_pipeline_run_noalloc on the GPU: the no-allocation entry point gives the same
group keys and int64 sums, bit for bit, as svm_pipeline_run on the same pipeline,
on two calls in a row (the second over a smaller input), into caller device
buffers; it refuses, without allocating, a pipeline shape whose cached state no
earlier run built, an output buffer too small for the result and a select pipeline."""
import numpy as np
import pytest

from shadowvm_amd import _lib
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns

pytestmark = pytest.mark.gpu


def _out_table(id, K, M, rows):
    if P.table_exists(id):
        P.table_drop(id)
    P.table_create(id, [P.INT32] * K + [P.INT64] * M, [4] * K + [8] * M, rows)
    return [P.table_column_device(id, c) for c in range(K + M)]


def _read(id, K, M, rows):
    return [np.asarray(P.table_column(id, c))[:rows] for c in range(K + M)]


def test_noalloc_equals_pipeline_run(engine):
    wl = ssb.four_way(seed=61, fact_rows=4_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000, prefix="na")
    ssb.load_device(wl)
    ssb.run_dims(wl)
    fq, src, refs = wl.fact
    P.table_view("na_small", src, 17, 1_500_001)
    # the existing entry point (which also builds the cached state and kernels)
    full = fq.Run(src, refs)
    small = fq.Run("na_small", refs)
    assert 0 < small.rows < full.rows
    cap = full.rows + 5
    ptrs = _out_table("na_out", 3, 1, cap)
    print(f"[noalloc] four-way: {full.rows} and {small.rows} groups", flush=True)
    n1 = P.run_noalloc(fq.plan(src, refs), ptrs, cap)
    assert n1 == full.rows
    assert_same_columns(_read("na_out", 3, 1, n1), full.columns)
    n2 = P.run_noalloc(fq.plan("na_small", refs), ptrs, cap)
    assert n2 == small.rows
    assert_same_columns(_read("na_out", 3, 1, n2), small.columns)
    # too small an output: refused, with the needed row count
    with pytest.raises(_lib.SvmError) as e:
        P.run_noalloc(fq.plan(src, refs), ptrs[:4], 3)
    assert e.value.code == _lib.ERR_INVALID
    # a select (dimension) pipeline is not a no-allocation run
    pt, dsrc, drefs = wl.dims[0]
    with pytest.raises(_lib.SvmError) as e:
        P.run_noalloc(pt.plan(dsrc, drefs), ptrs, cap)
    assert e.value.code == _lib.ERR_UNSUPPORTED
    P.table_drop("na_small")


def test_noalloc_scalar_and_unwarmed_shape(engine):
    wl = ssb.single_join(seed=62, fact_rows=2_000_003, prefix="nas")
    ssb.load_device(wl)
    ssb.run_dims(wl)
    fq, src, refs = wl.fact
    # a shape no earlier run compiled: refused without allocating anything
    ptrs = _out_table("nas_out", 0, 1, 1)
    with pytest.raises(_lib.SvmError) as e:
        P.run_noalloc(fq.plan(src, refs), ptrs, 1)
    assert e.value.code == _lib.ERR_STATE
    want = fq.Run(src, refs)
    P.table_view("nas_small", src, 5, 700_001)
    want_small = fq.Run("nas_small", refs)
    assert P.run_noalloc(fq.plan(src, refs), ptrs, 1) == 1
    assert_same_columns(_read("nas_out", 0, 1, 1), want.columns)
    assert P.run_noalloc(fq.plan("nas_small", refs), ptrs, 1) == 1
    assert_same_columns(_read("nas_out", 0, 1, 1), want_small.columns)
    assert want_small.columns[0][0] != want.columns[0][0]
    P.table_drop("nas_small")
